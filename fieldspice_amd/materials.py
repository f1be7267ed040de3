"""Per-cell material assembly: the source of the eps / sigma / kappa arrays.

Keeps the contract of ``fieldspice/materials.py`` that the field solve relies
on (``MaterialMap.eps()/sigma()`` return ABSOLUTE SI arrays of shape
``grid.shape_cells``, ``fieldspice/materials.py:1485-1500``), including the
order-independent effective-medium accumulation of partial fills
(``fieldspice/materials.py:1066-1141``):

    T(p_cell) = (1 - A) T(p_background) + sum_i f_i T(p_i)

with ``T`` the identity for the linear rule and ``1/x`` for the harmonic one,
``A`` the claimed coverage and ``f_i`` each phase's fill.  A phase that
over-fills a cell displaces earlier phases proportionally.

Only the electrical and thermal constants the elliptic solve needs are carried
(eps_r, sigma, kappa, rho*c_p); semiconductor transport parameters are not part
of this package.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .grid import RectilinearGrid
from .units import eps0

__all__ = ["Material", "LIBRARY", "get", "MaterialMap", "UnknownMaterial"]


class UnknownMaterial(KeyError, ValueError):
    """Unknown material name (both a KeyError and a ValueError)."""


@dataclass(frozen=True)
class Material:
    """Isotropic linear material (relative permittivity, S/m, W/(m K))."""

    name: str
    eps_r: float = 1.0
    sigma: float = 0.0
    kappa: float = 1.0
    rho_mass: float = 1000.0
    c_p: float = 1000.0
    kind: str = "dielectric"

    def __post_init__(self):
        if not self.name:
            raise ValueError("Material.name must be non-empty")
        if not np.isfinite(self.eps_r) or self.eps_r <= 0:
            raise ValueError(f"{self.name}: eps_r must be finite and > 0")
        if not np.isfinite(self.sigma) or self.sigma < 0:
            raise ValueError(f"{self.name}: sigma must be finite and >= 0")

    @property
    def eps(self) -> float:
        return self.eps_r * eps0

    def volumetric_heat_capacity(self) -> float:
        return self.rho_mass * self.c_p


# name: (eps_r, sigma [S/m], kappa [W/(m K)], rho [kg/m^3], c_p [J/(kg K)], kind)
# Values as listed by fieldspice/materials.py:682-860 for the materials an
# interconnect extraction uses.
_TABLE = {
    "vacuum": (1.0, 0.0, 1e-6, 1e-6, 1.0, "vacuum"),
    "air": (1.000589, 0.0, 0.0262, 1.204, 1005.0, "dielectric"),
    "si": (11.7, 2.83e-4, 148.0, 2329.0, 700.0, "semiconductor"),
    "sio2": (3.9, 1.0e-14, 1.4, 2200.0, 730.0, "dielectric"),
    "si3n4": (7.5, 1.0e-14, 30.0, 3170.0, 700.0, "dielectric"),
    "hfo2": (25.0, 1.0e-12, 1.1, 9680.0, 290.0, "dielectric"),
    "alumina": (9.8, 1.0e-12, 30.0, 3950.0, 880.0, "dielectric"),
    "aln": (8.9, 1.0e-11, 285.0, 3260.0, 740.0, "dielectric"),
    "fr4": (4.4, 1.0e-11, 0.30, 1850.0, 1200.0, "dielectric"),
    "poly": (11.7, 1.0e4, 30.0, 2320.0, 700.0, "conductor"),
    "cu": (1.0, 5.8e7, 401.0, 8960.0, 385.0, "conductor"),
    "al": (1.0, 3.77e7, 237.0, 2700.0, 897.0, "conductor"),
    "w": (1.0, 1.79e7, 173.0, 19250.0, 134.0, "conductor"),
}
LIBRARY: dict[str, Material] = {
    k: Material(k, eps_r=v[0], sigma=v[1], kappa=v[2], rho_mass=v[3],
                c_p=v[4], kind=v[5]) for k, v in _TABLE.items()}

_ALIASES = {"oxide": "sio2", "silicon_dioxide": "sio2", "nitride": "si3n4",
            "copper": "cu", "aluminum": "al", "aluminium": "al",
            "tungsten": "w", "silicon": "si", "vac": "vacuum"}


def get(name) -> Material:
    """Look up a material by (case-insensitive) name, or pass one through."""
    if isinstance(name, Material):
        return name
    if not isinstance(name, str):
        raise ValueError(f"material must be a str or Material, got {type(name).__name__}")
    key = name.strip().lower().replace("-", "_").replace(" ", "_")
    key = _ALIASES.get(key, key)
    if key not in LIBRARY:
        raise UnknownMaterial(f"unknown material {name!r}; known: {sorted(LIBRARY)}")
    return LIBRARY[key]


_LIN = ("linear", "parallel", "arithmetic")
_HAR = ("harmonic", "series")
_ATTR = {"eps": "eps_r", "sigma": "sigma", "kappa": "kappa",
         "rho_cp": "volumetric_heat_capacity"}


def _value(mat: Material, what: str) -> float:
    v = getattr(mat, _ATTR[what])
    return float(v() if callable(v) else v)


class MaterialMap:
    """Per-cell material properties on a grid.

    ``assign(mask, material, mix=..., mix_sigma=...)`` places a material into
    the cells of a boolean mask or a fill-fraction array; ``eps()`` and
    ``sigma()`` return absolute SI arrays for the solvers.
    """

    def __init__(self, grid: RectilinearGrid, background="vacuum"):
        if not isinstance(grid, RectilinearGrid):
            raise ValueError(f"grid must be a RectilinearGrid, got {type(grid).__name__}")
        self.grid = grid
        shape = grid.shape_cells
        self._materials = [get(background)]
        self._ids = np.zeros(shape, dtype=np.int32)
        self._cov = np.zeros(shape)
        self._acc = {w: np.zeros(shape) for w in _ATTR}
        self._dom = {w: np.zeros(shape, dtype=np.uint8) for w in _ATTR}
        self._blended = np.zeros(shape, dtype=bool)

    @property
    def background(self) -> Material:
        return self._materials[0]

    def _as_fill(self, mask) -> np.ndarray:
        a = np.asarray(mask)
        a = a.astype(float) if a.dtype == bool else a.astype(float, copy=False)
        if not np.all(np.isfinite(a)):
            raise ValueError("fill fractions must be finite")
        if a.size and (a.min() < 0.0 or a.max() > 1.0):
            raise ValueError("fill fractions must lie in [0, 1]")
        shape = self.grid.shape_cells
        if a.ndim == 0:
            return np.full(shape, float(a))
        if a.ndim != 3 or any(n not in (1, s) for n, s in zip(a.shape, shape)):
            raise ValueError(f"mask shape {a.shape} does not match the cell grid {shape}")
        return np.broadcast_to(a, shape).astype(float)

    def assign(self, mask, material, mix: str = "linear",
               mix_sigma: str = "linear") -> None:
        """Fill ``mask`` with ``material`` (order-independent accumulation).

        A cell that already holds a blend of two phases keeps its mixing rule;
        a single-phase cell may switch rule, its accumulator moving to the new
        domain by a reciprocal (``fieldspice/materials.py:1412-1449``).
        """
        mat = get(material)
        f = self._as_fill(mask)
        rules = {"eps": mix, "sigma": mix_sigma, "kappa": "linear",
                 "rho_cp": "linear"}
        touched = f > 0.0
        free = np.clip(1.0 - self._cov, 0.0, 1.0)
        over = f > free
        # displaced share of earlier phases when the new fill overruns the cell
        scale = np.where(over, (1.0 - f) / np.where(self._cov > 0, self._cov, 1.0), 1.0)
        new_acc = {}
        for what, rule in rules.items():
            if rule not in _LIN + _HAR:
                raise ValueError(f"unknown mixing rule {rule!r} for {what}")
            code = 1 if rule in _LIN else 2
            v = _value(mat, what)
            acc = self._acc[what]
            stale = (self._dom[what] != 0) & (self._dom[what] != code) & touched
            if np.any(stale & self._blended & (f < 1.0)):
                raise ValueError(f"cells already hold a mixture of {what} under "
                                 f"the other rule; cannot blend with {rule!r}")
            conv = stale & (f < 1.0)
            if np.any(conv):
                if np.any(acc[conv] <= 0.0):
                    raise ValueError(f"cannot switch the {what} rule in a cell "
                                     "holding a zero-valued phase")
                acc = acc.copy()
                acc[conv] = 1.0 / acc[conv]
            if code == 2:
                if v <= 0.0 or _value(self._materials[0], what) <= 0.0:
                    raise ValueError(f"harmonic mixing of {what} needs positive values")
                new_acc[what] = scale * acc + f / v
            else:
                new_acc[what] = scale * acc + f * v
        if mat not in self._materials:
            self._materials.append(mat)
        if not np.any(touched):
            return
        mid = self._materials.index(mat)
        self._acc = new_acc
        self._cov = np.clip(self._cov + f, 0.0, 1.0)
        self._ids = np.where(f >= 0.5, np.int32(mid), self._ids)
        self._blended[touched & (f < 1.0)] = True
        self._blended[f >= 1.0] = False
        for what, rule in rules.items():
            self._dom[what][touched] = 1 if rule in _LIN else 2

    def _effective(self, what: str) -> np.ndarray:
        bg = _value(self._materials[0], what)
        free = np.clip(1.0 - self._cov, 0.0, 1.0)
        out = free * bg + self._acc[what]
        har = self._dom[what] == 2
        if np.any(har):
            out[har] = 1.0 / (free[har] / bg + self._acc[what][har])
        return out

    def eps(self) -> np.ndarray:
        """Absolute permittivity per cell [F/m]."""
        return eps0 * self._effective("eps")

    def eps_r(self) -> np.ndarray:
        return self._effective("eps")

    def sigma(self) -> np.ndarray:
        """Conductivity per cell [S/m]."""
        return self._effective("sigma")

    def kappa(self) -> np.ndarray:
        """Thermal conductivity per cell [W/(m K)]."""
        return self._effective("kappa")

    def rho_cp(self) -> np.ndarray:
        """Volumetric heat capacity per cell [J/(m^3 K)]."""
        return self._effective("rho_cp")

    def coverage(self) -> np.ndarray:
        return self._cov.copy()

    @property
    def ids(self) -> np.ndarray:
        return self._ids.copy()

    def mask(self, material) -> np.ndarray:
        mat = get(material)
        if mat not in self._materials:
            return np.zeros(self.grid.shape_cells, dtype=bool)
        return self._ids == self._materials.index(mat)
