"""Shared pieces of the GPU time loops: Dirichlet groups and source bases."""

from __future__ import annotations

import numpy as np

from ..boundaries import WALLS, wall_nodes

__all__ = ["DirichletGroups", "source_basis"]


class DirichletGroups:
    """Fixed nodes as groups sharing one value per time point.

    Terminals and Dirichlet walls are claimed in the given order; a later
    claim wins a shared node, which reproduces fieldspice's "last entry wins"
    deduplication (``fieldspice/solvers/eqs.py:108-129``: walls after
    terminals; ``fieldspice/boundaries.py:963-1029``: later wall wins).
    """

    def __init__(self, grid, terminals=(), bc=None, walls_last: bool = True):
        self.grid = grid
        self.owner = np.full(grid.n_nodes, -1, dtype=np.int32)
        self.sources: list = []

        def claim_terms():
            for t in terminals:
                if t.driven != "voltage":
                    continue
                self.owner[t.nodes] = len(self.sources)
                self.sources.append(("terminal", t))

        def claim_walls():
            if bc is None:
                return
            for w in WALLS:
                wb = getattr(bc, w)
                if not wb.is_dirichlet_like:
                    continue
                self.owner[wall_nodes(grid, w)] = len(self.sources)
                self.sources.append(("wall", wb))

        if walls_last:
            claim_terms()
            claim_walls()
        else:
            claim_walls()
            claim_terms()

    @property
    def ngroups(self) -> int:
        return len(self.sources)

    @property
    def fixed(self) -> np.ndarray:
        return np.flatnonzero(self.owner >= 0)

    def values_at(self, t: float) -> np.ndarray:
        out = np.zeros(len(self.sources))
        for g, (kind, src) in enumerate(self.sources):
            if kind == "terminal":
                out[g] = float(src.value_at(t))
            else:
                v = np.asarray(src.value_at(t), dtype=float)
                if v.ndim != 0:
                    raise NotImplementedError(
                        "spatially varying Dirichlet walls are not supported by the GPU "
                        "time loop; split the wall into terminals instead")
                out[g] = float(v)
        return out

    def node_values(self, t: float) -> np.ndarray:
        """(fixed node indices, their values) at time t."""
        vals = self.values_at(t)
        fixed = self.fixed
        return fixed, vals[self.owner[fixed]] if fixed.size else np.empty(0)


def source_basis(vectors, max_basis: int = 8):
    """Express a sequence of node vectors as ``scale[k] @ basis``.

    Vectors that are scalar multiples of an earlier basis vector reuse it, so a
    constant source or a source switched on and off costs one basis vector.
    Returns (basis (m, n), scales (len(vectors), m)).
    """
    basis: list[np.ndarray] = []
    norms: list[float] = []
    scales = np.zeros((len(vectors), max_basis))
    for k, v in enumerate(vectors):
        v = np.asarray(v, dtype=float).ravel()
        nv = float(np.linalg.norm(v))
        if nv == 0.0:
            continue
        placed = False
        for m, (b, nb) in enumerate(zip(basis, norms)):
            c = float(v @ b) / (nb * nb)
            if np.array_equal(v, c * b) or np.allclose(v, c * b, rtol=1e-15, atol=0.0):
                scales[k, m] = c
                placed = True
                break
        if not placed:
            if len(basis) == max_basis:
                raise NotImplementedError(
                    f"the source takes more than {max_basis} independent shapes over the run; "
                    "the GPU time loop supports sources that are combinations of a few fixed "
                    "node vectors")
            basis.append(v)
            norms.append(nv)
            scales[k, len(basis) - 1] = 1.0
    m = max(len(basis), 0)
    return (np.array(basis) if basis else np.zeros((0, 0))), scales[:, :m]
