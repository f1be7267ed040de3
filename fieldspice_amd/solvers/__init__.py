"""GPU solvers: the elliptic field solve behind capacitance extraction and the
EQS / thermal implicit steps."""

from .base import ConvergenceError, Result, SolverConfig, Terminal
from .linear import GridLinearSystem
from .poisson import NonlinearPoissonSolver, PoissonSolver, boltzmann_equilibrium_charge
from .eqs import EQSSolver
from .thermal import ThermalSolver, joule_heating_nodes
from .electrothermal import ElectroThermalSolver, ThermalRunaway, runaway_threshold
from . import checks

__all__ = ["ConvergenceError", "Result", "SolverConfig", "Terminal", "GridLinearSystem",
           "PoissonSolver", "NonlinearPoissonSolver", "boltzmann_equilibrium_charge",
           "EQSSolver", "ThermalSolver", "joule_heating_nodes", "ElectroThermalSolver",
           "ThermalRunaway", "runaway_threshold", "checks"]
