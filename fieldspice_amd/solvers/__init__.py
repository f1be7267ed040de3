"""GPU solvers: the elliptic field solve behind capacitance extraction and the
EQS / thermal implicit steps."""

from .base import ConvergenceError, Result, SolverConfig, Terminal
from .linear import GridLinearSystem
from .poisson import PoissonSolver

__all__ = ["ConvergenceError", "Result", "SolverConfig", "Terminal", "GridLinearSystem",
           "PoissonSolver"]
