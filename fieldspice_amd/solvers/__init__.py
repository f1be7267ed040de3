"""GPU solvers: the elliptic field solve behind capacitance extraction and the
EQS / thermal implicit steps."""

from .base import ConvergenceError, Result, SolverConfig, Terminal
from .linear import GridLinearSystem
from .poisson import PoissonSolver
from .eqs import EQSSolver
from .thermal import ThermalSolver, joule_heating_nodes

__all__ = ["ConvergenceError", "Result", "SolverConfig", "Terminal", "GridLinearSystem",
           "PoissonSolver", "EQSSolver", "ThermalSolver", "joule_heating_nodes"]
