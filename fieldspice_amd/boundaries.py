"""Wall boundary conditions as data (``fieldspice/boundaries.py``).

Semantics kept from the reference:

* Homogeneous Neumann is the natural condition of ``G^T M G`` and needs no
  action (``fieldspice/boundaries.py:13-26``).
* Dirichlet beats Neumann; among Dirichlet walls the later one in
  ``xlo, xhi, ylo, yhi, zlo, zhi`` wins at shared nodes
  (``fieldspice/boundaries.py:28-46,963-1029``).
* ``wall_nodes`` returns ascending flat C-order (z-fastest) indices and
  ``wall_dual_areas`` the matching exposed dual-face areas
  (``fieldspice/boundaries.py:154-218``).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, ClassVar

import numpy as np

from .grid import RectilinearGrid

__all__ = ["WALLS", "WALL_AXIS", "WALL_SIDE", "BC", "Dirichlet", "Neumann",
           "Periodic", "Symmetry", "BoundarySpec", "wall_nodes", "wall_dual_areas"]

WALLS = ("xlo", "xhi", "ylo", "yhi", "zlo", "zhi")
WALL_AXIS = {"xlo": 0, "xhi": 0, "ylo": 1, "yhi": 1, "zlo": 2, "zhi": 2}
WALL_SIDE = {"xlo": 0, "xhi": 1, "ylo": 0, "yhi": 1, "zlo": 0, "zhi": 1}


def _check_wall(wall: str) -> str:
    if wall not in WALL_AXIS:
        raise ValueError(f"unknown wall {wall!r}; expected one of {WALLS}")
    return wall


def wall_nodes(grid: RectilinearGrid, wall: str) -> np.ndarray:
    """Ascending flat node indices on one wall."""
    _check_wall(wall)
    shape = grid.shape_nodes
    stride = (shape[1] * shape[2], shape[2], 1)
    ax = WALL_AXIS[wall]
    plane = 0 if WALL_SIDE[wall] == 0 else shape[ax] - 1
    d0, d1 = (d for d in range(3) if d != ax)
    return (plane * stride[ax]
            + (np.arange(shape[d0], dtype=np.intp) * stride[d0])[:, None]
            + (np.arange(shape[d1], dtype=np.intp) * stride[d1])[None, :]).ravel()


def wall_dual_areas(grid: RectilinearGrid, wall: str) -> np.ndarray:
    """Exposed dual-box face area of each wall node, ordered as ``wall_nodes``."""
    _check_wall(wall)
    duals = (grid.hxd, grid.hyd, grid.hzd)
    d0, d1 = (d for d in range(3) if d != WALL_AXIS[wall])
    return np.multiply.outer(duals[d0], duals[d1]).ravel()


class BC:
    """A condition on one wall."""

    kind: ClassVar[str] = "bc"
    is_dirichlet_like = False
    is_natural = False

    def prescribed_value(self, t: float = 0.0):
        return None


@dataclass(frozen=True)
class Dirichlet(BC):
    """Prescribed potential: a constant, or ``f(t)`` returning a scalar or a
    per-wall-node array."""

    value: float | Callable[[float], float] = 0.0
    kind: ClassVar[str] = "dirichlet"
    is_dirichlet_like = True

    def __post_init__(self):
        if not callable(self.value):
            try:
                float(self.value)
            except (TypeError, ValueError) as exc:
                raise ValueError("Dirichlet value must be a float or a callable") from exc

    @property
    def time_dependent(self) -> bool:
        return callable(self.value)

    def value_at(self, t: float = 0.0) -> np.ndarray:
        v = self.value(float(t)) if callable(self.value) else self.value
        arr = np.asarray(v, dtype=float)
        if arr.ndim > 1:
            raise ValueError("Dirichlet value must be scalar or 1-D")
        if not np.all(np.isfinite(arr)):
            raise ValueError(f"Dirichlet value is not finite at t = {t} s")
        return arr

    prescribed_value = value_at


@dataclass(frozen=True)
class Neumann(BC):
    """Prescribed outward normal flux density; 0 is the natural condition."""

    flux: float | Callable[[float], float] = 0.0
    kind: ClassVar[str] = "neumann"

    @property
    def is_natural(self) -> bool:  # type: ignore[override]
        return not callable(self.flux) and float(self.flux) == 0.0

    def flux_at(self, t: float = 0.0) -> np.ndarray:
        f = self.flux(float(t)) if callable(self.flux) else self.flux
        return np.asarray(f, dtype=float)


@dataclass(frozen=True)
class Symmetry(BC):
    """A mirror plane (``fieldspice/boundaries.py:527-605``): ``"pmc"`` (even
    potential, ``dphi/dn = 0``: natural, as Neumann) or ``"pec"`` (odd
    potential, constant on the plane: Dirichlet at ``value``)."""

    wall: str = "pmc"
    value: float = 0.0
    kind: ClassVar[str] = "symmetry"

    def __post_init__(self):
        kind = str(self.wall).lower()
        if kind not in ("pec", "pmc"):
            raise ValueError(f"Symmetry wall must be 'pec' or 'pmc', got {self.wall!r}")
        object.__setattr__(self, "wall", kind)
        try:
            float(self.value)
        except (TypeError, ValueError) as exc:
            raise ValueError("Symmetry value must be a float [V]") from exc

    @property
    def is_electric(self) -> bool:
        return self.wall == "pec"

    @property
    def is_dirichlet_like(self) -> bool:  # type: ignore[override]
        return self.is_electric

    @property
    def is_natural(self) -> bool:  # type: ignore[override]
        return not self.is_electric

    def value_at(self, t: float = 0.0) -> np.ndarray:
        if not self.is_electric:
            raise ValueError("a PMC symmetry plane prescribes no potential")
        return np.asarray(float(self.value))

    def prescribed_value(self, t: float = 0.0):
        return self.value_at(t) if self.is_electric else None

    def as_scalar_bc(self) -> BC:
        """The scalar-potential condition it stands for."""
        return Dirichlet(float(self.value)) if self.is_electric else Neumann(0.0)


@dataclass(frozen=True)
class Periodic(BC):
    """Periodic wall pair.  Accepted as data; the GPU solvers reject it."""

    offset: float = 0.0
    phase: float = 0.0
    kind: ClassVar[str] = "periodic"


@dataclass(frozen=True)
class BoundarySpec:
    """Conditions on the six walls; default all homogeneous Neumann."""

    xlo: BC = Neumann()
    xhi: BC = Neumann()
    ylo: BC = Neumann()
    yhi: BC = Neumann()
    zlo: BC = Neumann()
    zhi: BC = Neumann()

    def __post_init__(self):
        for w in WALLS:
            if not isinstance(getattr(self, w), BC):
                raise ValueError(f"boundary {w} must be a BC instance")

    @classmethod
    def all_neumann(cls, grid=None) -> "BoundarySpec":
        return cls()

    @classmethod
    def all_dirichlet(cls, v=0.0, grid: RectilinearGrid | None = None) -> "BoundarySpec":
        """Dirichlet on every wall; with ``grid``, collapsed axes stay Neumann."""
        walls = {}
        for w in WALLS:
            collapsed = grid is not None and grid.ncell[WALL_AXIS[w]] == 1
            walls[w] = Neumann() if collapsed else Dirichlet(v)
        return cls(**walls)

    def get(self, wall: str) -> BC:
        return getattr(self, _check_wall(wall))

    __getitem__ = get

    def items(self):
        return tuple((w, getattr(self, w)) for w in WALLS)

    @property
    def is_time_dependent(self) -> bool:
        return any(getattr(bc, "time_dependent", False) for _, bc in self.items())

    def kinds(self) -> tuple[str, ...]:
        return tuple(bc.kind for _, bc in self.items())

    def validate(self, grid=None) -> None:
        for a in range(3):
            lo, hi = getattr(self, WALLS[2 * a]), getattr(self, WALLS[2 * a + 1])
            if isinstance(lo, Periodic) != isinstance(hi, Periodic):
                raise ValueError(f"axis {'xyz'[a]}: Periodic must be set on both walls")

    def require(self, allowed, context: str = "this solver") -> None:
        bad = [(w, bc.kind) for w, bc in self.items() if bc.kind not in allowed]
        if bad:
            listed = ", ".join(f"{w}={k}" for w, k in bad)
            raise ValueError(f"{context} does not support boundary condition(s): "
                             f"{listed}. Supported kinds: {sorted(allowed)}")

    def dirichlet_nodes(self, grid: RectilinearGrid, t: float = 0.0):
        """(ascending unique node indices, values); later wall wins."""
        idx_parts, val_parts = [], []
        for w in WALLS:
            bc = getattr(self, w)
            if not bc.is_dirichlet_like:
                continue
            idx = wall_nodes(grid, w)
            v = bc.prescribed_value(t)
            try:
                vals = np.broadcast_to(v, idx.shape).astype(float, copy=True)
            except ValueError as exc:
                raise ValueError(f"wall {w}: Dirichlet value of shape {np.shape(v)} "
                                 f"cannot be broadcast to {idx.size} nodes") from exc
            idx_parts.append(idx)
            val_parts.append(vals)
        if not idx_parts:
            return np.empty(0, dtype=np.intp), np.empty(0)
        return unique_last(np.concatenate(idx_parts), np.concatenate(val_parts))

    def dirichlet_walls(self, grid: RectilinearGrid):
        """Per-node owning wall ordinal (-1 where no Dirichlet wall), later wins.

        Used by the device time loops, which evaluate one value per wall per
        step instead of a value per node.
        """
        owner = np.full(grid.n_nodes, -1, dtype=np.int32)
        for n, w in enumerate(WALLS):
            if getattr(self, w).is_dirichlet_like:
                owner[wall_nodes(grid, w)] = n
        return owner

    def neumann_load(self, grid: RectilinearGrid, t: float = 0.0) -> np.ndarray:
        """RHS term ``-flux*A`` of driven Neumann walls, zero on Dirichlet nodes."""
        load = np.zeros(grid.n_nodes)
        for w in WALLS:
            bc = getattr(self, w)
            if not isinstance(bc, Neumann) or bc.is_natural:
                continue
            idx = wall_nodes(grid, w)
            load[idx] += -np.broadcast_to(bc.flux_at(t), idx.shape) * wall_dual_areas(grid, w)
        fixed, _ = self.dirichlet_nodes(grid, t)
        load[fixed] = 0.0
        return load


def unique_last(idx: np.ndarray, val: np.ndarray):
    """Deduplicate indices keeping the LAST occurrence's value
    (``fieldspice/solvers/poisson.py:722-727``)."""
    idx = np.asarray(idx)
    val = np.asarray(val, dtype=float)
    if idx.size == 0:
        return idx.astype(np.intp), val
    uniq, first_rev = np.unique(idx[::-1], return_index=True)
    return uniq.astype(np.intp), val[::-1][first_rev]
