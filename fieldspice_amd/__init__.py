"""fieldspice_amd -- fieldspice's elliptic field solve on AMD MI355X (gfx950).

The public names match fieldspice's (``RectilinearGrid``, ``MaterialMap``,
``Terminal``, ``SolverConfig``, ``extraction.capacitance_matrix``, the Poisson,
EQS and thermal solvers); every solve runs on the GPU through the C ABI of
``libfieldspice_hip.so`` (include/fieldspice_hip.h).  There is no CPU solver.
"""

from . import boundaries, extraction, physics, units
from .boundaries import BoundarySpec, Dirichlet, Neumann
from .grid import RectilinearGrid, graded_1d
from .materials import MaterialMap
from .solvers.base import ConvergenceError, Result, SolverConfig, Terminal
from .solvers.poisson import NonlinearPoissonSolver, PoissonSolver, boltzmann_equilibrium_charge
from .solvers.eqs import EQSSolver
from .solvers.thermal import ThermalSolver, joule_heating_nodes
from .solvers.electrothermal import ElectroThermalSolver, ThermalRunaway, runaway_threshold
from .physics import PhysicsOptions

__version__ = "0.1.0"

__all__ = ["RectilinearGrid", "graded_1d", "MaterialMap", "BoundarySpec", "Dirichlet",
           "Neumann", "Terminal", "SolverConfig", "Result", "ConvergenceError",
           "PoissonSolver", "NonlinearPoissonSolver", "boltzmann_equilibrium_charge", "EQSSolver", "ThermalSolver", "joule_heating_nodes", "extraction",
           "units", "boundaries", "physics", "PhysicsOptions", "ElectroThermalSolver",
           "ThermalRunaway", "runaway_threshold"]
