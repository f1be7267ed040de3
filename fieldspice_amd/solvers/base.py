"""Solver contract types: SolverConfig, Terminal, Result, ConvergenceError.

Same fields and semantics as ``fieldspice/solvers/base.py:35-202``, with the
``linear_solver`` vocabulary extended by the GPU methods of this package.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Any, Callable, Sequence

import numpy as np

__all__ = ["ConvergenceError", "SolverConfig", "Terminal", "Result", "LINEAR_SOLVERS"]

# "auto", "direct", "cg" and "amg" are fieldspice's names; all of them run the
# GPU Jacobi-preconditioned CG here (the direct and AMG paths have no GPU
# counterpart and their answers agree to the solve tolerance).
LINEAR_SOLVERS = ("auto", "direct", "cg", "amg", "hip_pcg", "hip_mixed", "hip_fcg")


class ConvergenceError(RuntimeError):
    """Iterative solve failed; carries the residual history
    (``fieldspice/solvers/base.py:35-48``)."""

    def __init__(self, msg: str, history: Sequence[float] = (),
                 last_state: np.ndarray | None = None):
        super().__init__(msg)
        self.history = list(history)
        self.last_state = last_state


@dataclass
class SolverConfig:
    """Knobs shared by the solvers (``fieldspice/solvers/base.py:51-82``).

    GPU additions:

    linear_solver
        ``"hip_pcg"`` (FP64 CG on the equilibrated system, the reference's
        algorithm), ``"hip_mixed"`` (FP32 CG cycles with FP64 residual
        refresh), ``"hip_fcg"`` (flexible CG with the learned preconditioner).
    inner_tol
        Residual reduction of one FP32 cycle before the FP64 refresh
        (``None``: 1e-4 for ``hip_mixed``, 1e-3 for ``hip_fcg``).
    preconditioner
        ``hip_fcg`` only: a ``precond.LearnedPreconditioner``; ``None`` uses
        the trained network shipped with the package.
    maxiter
        Krylov cap; ``None`` is fieldspice's ``max(500, 10*sqrt(n))``.
    deflate_islands
        ``hip_pcg``/``hip_mixed``: add one coarse vector per floating
        high-conductance component ("island": a copper box in SiO2 that touches
        no terminal) to the preconditioner.  Islands are the components of the
        edges stronger than ``sqrt(w_min*w_max)`` when ``w_max/w_min`` exceeds
        ``island_contrast``; without them such systems stall Jacobi-CG (and
        FP32 cycles cannot resolve them at all).
    """
    linear_solver: str = "auto"
    tol: float = 1e-10
    newton_tol: float = 1e-8
    max_newton: int = 60
    damping: float = 1.0
    verbose: int = 0
    store_every: int = 1
    dtype: type = np.float64
    inner_tol: float | None = None
    maxiter: int | None = None
    check_every: int = 0
    preconditioner: object = None
    deflate_islands: bool = False
    island_contrast: float = 1e3

    def __post_init__(self):
        if self.linear_solver not in LINEAR_SOLVERS:
            raise ValueError(f"unknown linear_solver {self.linear_solver!r}; expected one "
                             f"of {LINEAR_SOLVERS}")


@dataclass
class Terminal:
    """An electrode: a set of nodes at a common potential
    (``fieldspice/solvers/base.py:85-133``)."""
    name: str
    nodes: np.ndarray
    voltage: float | Callable[[float], float] | None = None
    current: float | Callable[[float], float] | None = None
    circuit_node: str | None = None

    def __post_init__(self):
        self.nodes = np.unique(np.asarray(self.nodes, dtype=np.intp).ravel())
        if self.voltage is not None and self.current is not None:
            raise ValueError(f"terminal {self.name!r}: set voltage or current, not both")

    def value_at(self, t: float) -> float | None:
        src = self.voltage if self.voltage is not None else self.current
        if src is None:
            return None
        return float(src(t)) if callable(src) else float(src)

    @property
    def driven(self) -> str:
        if self.voltage is not None:
            return "voltage"
        if self.current is not None:
            return "current"
        return "float"


@dataclass
class Result:
    """Uniform output container (``fieldspice/solvers/base.py:136-202``)."""
    grid: Any
    t: np.ndarray = field(default_factory=lambda: np.zeros(0))
    fields: dict[str, np.ndarray] = field(default_factory=dict)
    terminals: dict[str, dict[str, np.ndarray]] = field(default_factory=dict)
    scalars: dict[str, np.ndarray] = field(default_factory=dict)
    meta: dict[str, Any] = field(default_factory=dict)

    def v(self, terminal: str) -> np.ndarray:
        return self.terminals[terminal]["v"]

    def i(self, terminal: str) -> np.ndarray:
        return self.terminals[terminal]["i"]

    def field(self, name: str, step: int = -1) -> np.ndarray:
        return self.fields[name][step]

    def final(self, name: str) -> np.ndarray:
        return self.fields[name][-1]

    def save(self, path: str) -> None:
        """``.npz`` dump of fields, scalars and terminal traces."""
        arrays = {f"field_{k}": v for k, v in self.fields.items()}
        arrays.update({f"scalar_{k}": v for k, v in self.scalars.items()})
        for name, tr in self.terminals.items():
            for q, v in tr.items():
                arrays[f"term_{name}_{q}"] = v
        np.savez(path, t=self.t, **arrays)


class _Timer:
    def __init__(self):
        self.t0 = time.perf_counter()

    def finish(self, res: Result, name: str, assumptions, grid, **extra) -> Result:
        res.meta.setdefault("solver", name)
        res.meta.setdefault("assumptions", list(assumptions))
        res.meta.setdefault("wall_time", time.perf_counter() - self.t0)
        res.meta.setdefault("grid_cells", grid.ncell)
        res.meta.update(extra)
        return res


def check_terminals(terminals, grid) -> tuple[Terminal, ...]:
    """Validation of ``fieldspice/solvers/poisson.py:759-787``: names unique,
    node sets non-empty, in range and pairwise disjoint (vectorised)."""
    if terminals is None:
        return ()
    if isinstance(terminals, Terminal):
        raise ValueError("pass a sequence of Terminal, not a bare Terminal")
    out = tuple(terminals)
    names: set[str] = set()
    owner = None
    for t in out:
        if not isinstance(t, Terminal):
            raise ValueError(f"expected Terminal instances, got {type(t).__name__}")
        if t.name in names:
            raise ValueError(f"duplicate terminal name {t.name!r}")
        names.add(t.name)
        if t.nodes.size == 0:
            raise ValueError(f"terminal {t.name!r} has no nodes")
        if t.nodes.min() < 0 or t.nodes.max() >= grid.n_nodes:
            raise ValueError(f"terminal {t.name!r} references node index outside "
                             f"[0, {grid.n_nodes})")
    if out:
        owner = np.full(grid.n_nodes, -1, dtype=np.int32)
        for k, t in enumerate(out):
            clash = owner[t.nodes] >= 0
            if np.any(clash):
                nd = int(t.nodes[np.argmax(clash)])
                raise ValueError(f"terminals {out[owner[nd]].name!r} and {t.name!r} share "
                                 f"node {nd}; electrodes must be disjoint")
            owner[t.nodes] = k
    return out
