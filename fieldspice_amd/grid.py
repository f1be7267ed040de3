"""Tensor-product rectilinear grid: the geometry every operator is built on.

Follows the conventions of ``fieldspice/grid.py:221-546``:

* ``(Nx, Ny, Nz)`` cells, ``(Nx+1, Ny+1, Nz+1)`` nodes.  A collapsed axis is a
  single cell of user-given thickness (default 1 m), so it still has two node
  planes.
* Primal widths ``h = diff(nodes)``; dual (node-centred) widths ``hd`` are the
  mean of the two flanking cells, and half a cell at either end
  (``fieldspice/grid.py:234-238``).
* Host arrays are C-ordered with z fastest: the flat node index of ``(i, j, k)``
  is ``(i*(Ny+1) + j)*(Nz+1) + k`` (``fieldspice/grid.py:441-444``).  The GPU
  library stores nodes x-fastest and converts at its boundary.
"""

from __future__ import annotations

from typing import Sequence

import numpy as np

__all__ = ["RectilinearGrid", "graded_1d"]


def graded_1d(a: float, b: float, n: int, ratio: float = 1.0,
              reverse: bool = False) -> np.ndarray:
    """``n+1`` node coordinates on ``[a, b]`` whose widths grow by ``ratio``.

    Same contract as ``fieldspice/grid.py:67-86``.
    """
    if n < 1:
        raise ValueError("n must be >= 1")
    if ratio <= 0:
        raise ValueError("ratio must be positive")
    widths = np.ones(n) if abs(ratio - 1.0) < 1e-12 else ratio ** np.arange(n)
    if reverse:
        widths = widths[::-1]
    widths = widths * ((b - a) / widths.sum())
    return np.concatenate([[a], a + np.cumsum(widths)])


class _Axis:
    __slots__ = ("nodes", "h", "hd", "collapsed")

    def __init__(self, coords, thickness: float):
        if coords is None:
            nodes = np.array([0.0, float(thickness)])
            collapsed = True
        else:
            nodes = np.asarray(coords, dtype=float).ravel()
            if nodes.size < 2:
                raise ValueError("an axis needs at least 2 node coordinates")
            if np.any(np.diff(nodes) <= 0):
                raise ValueError("node coordinates must be strictly increasing")
            collapsed = nodes.size == 2
        h = np.diff(nodes)
        hd = np.empty(nodes.size)
        hd[0] = 0.5 * h[0]
        hd[-1] = 0.5 * h[-1]
        hd[1:-1] = 0.5 * (h[:-1] + h[1:])
        self.nodes, self.h, self.hd, self.collapsed = nodes, h, hd, collapsed


class RectilinearGrid:
    """Graded axis-aligned grid; API of ``fieldspice/grid.py:242-546``.

    Parameters
    ----------
    x, y, z
        Node coordinates per axis; ``None`` collapses that axis to one cell of
        ``thickness[d]``.
    thickness
        Extent used for collapsed axes.
    """

    def __init__(self, x, y=None, z=None,
                 thickness: tuple[float, float, float] = (1.0, 1.0, 1.0)):
        self.ax = _Axis(x, thickness[0])
        self.ay = _Axis(y, thickness[1])
        self.az = _Axis(z, thickness[2])
        self._axes = (self.ax, self.ay, self.az)

    # sizes ---------------------------------------------------------------
    Nx = property(lambda self: self.ax.h.size)
    Ny = property(lambda self: self.ay.h.size)
    Nz = property(lambda self: self.az.h.size)

    @property
    def ncell(self) -> tuple[int, int, int]:
        return (self.Nx, self.Ny, self.Nz)

    shape_cells = ncell

    @property
    def shape_nodes(self) -> tuple[int, int, int]:
        return (self.Nx + 1, self.Ny + 1, self.Nz + 1)

    @property
    def shape_edges(self) -> tuple[tuple[int, int, int], ...]:
        nx, ny, nz = self.ncell
        return ((nx, ny + 1, nz + 1), (nx + 1, ny, nz + 1), (nx + 1, ny + 1, nz))

    @property
    def ndim_effective(self) -> int:
        return sum(0 if a.collapsed else 1 for a in self._axes) or 1

    @property
    def n_nodes(self) -> int:
        return int(np.prod(self.shape_nodes))

    @property
    def n_cells(self) -> int:
        return int(np.prod(self.shape_cells))

    @property
    def n_edges_each(self) -> tuple[int, int, int]:
        return tuple(int(np.prod(s)) for s in self.shape_edges)  # type: ignore

    @property
    def n_edges(self) -> int:
        return int(sum(self.n_edges_each))

    # coordinates and widths ----------------------------------------------
    xn = property(lambda self: self.ax.nodes)
    yn = property(lambda self: self.ay.nodes)
    zn = property(lambda self: self.az.nodes)
    hx = property(lambda self: self.ax.h)
    hy = property(lambda self: self.ay.h)
    hz = property(lambda self: self.az.h)
    hxd = property(lambda self: self.ax.hd)
    hyd = property(lambda self: self.ay.hd)
    hzd = property(lambda self: self.az.hd)

    @property
    def xc(self) -> np.ndarray:
        return 0.5 * (self.xn[:-1] + self.xn[1:])

    @property
    def yc(self) -> np.ndarray:
        return 0.5 * (self.yn[:-1] + self.yn[1:])

    @property
    def zc(self) -> np.ndarray:
        return 0.5 * (self.zn[:-1] + self.zn[1:])

    @property
    def bounds(self) -> tuple[tuple[float, float], ...]:
        return tuple((float(a.nodes[0]), float(a.nodes[-1])) for a in self._axes)

    # metric arrays ---------------------------------------------------------
    def cell_volumes(self) -> np.ndarray:
        return np.multiply.outer(np.multiply.outer(self.hx, self.hy), self.hz)

    def node_volumes(self) -> np.ndarray:
        """Dual box volumes, shape ``shape_nodes`` (``fieldspice/grid.py:388-396``)."""
        return np.multiply.outer(np.multiply.outer(self.hxd, self.hyd), self.hzd)

    def edge_lengths(self):
        """Primal edge lengths per family (``fieldspice/grid.py:398-405``)."""
        sx, sy, sz = self.shape_edges
        return (np.broadcast_to(self.hx[:, None, None], sx).copy(),
                np.broadcast_to(self.hy[None, :, None], sy).copy(),
                np.broadcast_to(self.hz[None, None, :], sz).copy())

    def edge_dual_areas(self):
        """Dual face area pierced by each edge (``fieldspice/grid.py:407-420``)."""
        sx, sy, sz = self.shape_edges
        return (np.broadcast_to(self.hyd[None, :, None] * self.hzd[None, None, :], sx).copy(),
                np.broadcast_to(self.hxd[:, None, None] * self.hzd[None, None, :], sy).copy(),
                np.broadcast_to(self.hxd[:, None, None] * self.hyd[None, :, None], sz).copy())

    # index helpers ---------------------------------------------------------
    def node_index(self, i, j, k) -> np.ndarray:
        return np.ravel_multi_index((i, j, k), self.shape_nodes)

    def node_ids(self) -> np.ndarray:
        """``arange(n_nodes)`` shaped ``shape_nodes`` (the usual ``nid`` idiom)."""
        return np.arange(self.n_nodes).reshape(self.shape_nodes)

    def node_coords(self):
        return np.meshgrid(self.xn, self.yn, self.zn, indexing="ij")

    def nearest_node(self, pt: Sequence[float]) -> tuple[int, int, int]:
        return tuple(int(np.argmin(np.abs(a.nodes - p)))
                     for a, p in zip(self._axes, pt))  # type: ignore

    def courant_dt(self, eps_r_min: float = 1.0, mu_r_min: float = 1.0) -> float:
        """Explicit-FDTD stability step (``fieldspice/grid.py:483-501``), used to
        report how many explicit steps one implicit step replaces."""
        from .units import c0
        v = 1.0 / np.sqrt(eps_r_min * mu_r_min)
        inv2 = sum(1.0 / float(a.h.min()) ** 2 for a in self._axes if not a.collapsed)
        if inv2 == 0.0:
            inv2 = 1.0 / float(self.ax.h.min()) ** 2
        return 1.0 / (c0 * v * np.sqrt(inv2))

    def __repr__(self) -> str:
        nx, ny, nz = self.ncell
        return f"<RectilinearGrid {nx}x{ny}x{nz}, {self.n_nodes:,} nodes>"

    @classmethod
    def uniform(cls, extent, n, thickness=(1.0, 1.0, 1.0)) -> "RectilinearGrid":
        """Uniform grid with ``n[d]`` cells over ``extent[d]``; missing axes collapse."""
        axes: list = [None, None, None]
        for d, (ext, ncell) in enumerate(zip(extent, n)):
            axes[d] = np.linspace(ext[0], ext[1], int(ncell) + 1)
        return cls(axes[0], axes[1], axes[2], thickness=thickness)
