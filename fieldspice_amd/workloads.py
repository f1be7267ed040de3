"""Seeded extraction geometries: the parallel plate and layered multi-conductor
problems the solver is measured on.

``layered_boxes`` builds a graded grid, a stack of dielectric layers whose
interfaces sit on node planes, and rectangular conductor boxes placed by
rejection sampling with at least one free cell between boxes (so terminal node
sets stay disjoint, as ``fieldspice/solvers/poisson.py:780-786`` requires).
Each box becomes a terminal owning every node of its closed box; its cells keep
the permittivity of the layer they sit in.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .grid import RectilinearGrid
from .solvers.base import Terminal
from .units import eps0

__all__ = ["Extraction", "parallel_plate", "layered_boxes", "c2_geometry", "multi_conductor",
           "batch_instances", "eqs_two_phase"]


@dataclass
class Extraction:
    grid: RectilinearGrid
    eps_cell: np.ndarray
    terminals: list
    boxes: list          # (i0, i1, j0, j1, k0, k1) cell ranges, half-open
    layer_eps_r: list
    seed: int


def parallel_plate(n=(24, 10, 8), extent=(2e-6, 5e-6, 4e-6), eps_r=3.9) -> Extraction:
    """Plates on the x=0 (1 V) and x=last (0 V) node planes of a uniform grid."""
    g = RectilinearGrid.uniform([(0, extent[0]), (0, extent[1]), (0, extent[2])], n)
    nid = g.node_ids()
    terms = [Terminal("hi", nid[0].ravel(), voltage=1.0),
             Terminal("lo", nid[-1].ravel(), voltage=0.0)]
    return Extraction(g, np.full(g.shape_cells, eps_r * eps0), terms, [], [eps_r], 0)


def layered_boxes(nodes=(192, 192, 96), h_xy=50e-9, z_first=20e-9, z_last=400e-9,
                  n_layers=3, eps_choices=(2.7, 3.9, 7.5), n_boxes=8,
                  side=(4, 40), seed=0, max_tries=100000) -> Extraction:
    rng = np.random.default_rng(seed)
    nx, ny, nz = nodes
    Nx, Ny, Nz = nx - 1, ny - 1, nz - 1
    x = np.arange(nx) * h_xy
    y = np.arange(ny) * h_xy
    if Nz > 1:
        ratio = (z_last / z_first) ** (1.0 / (Nz - 1))
        hz = z_first * ratio ** np.arange(Nz)
    else:
        hz = np.array([z_first])
    z = np.concatenate([[0.0], np.cumsum(hz)])
    g = RectilinearGrid(x, y, z)
    # layer interfaces on node planes (distinct, strictly inside)
    cuts = np.sort(rng.choice(np.arange(1, Nz), size=n_layers - 1, replace=False))
    bounds = np.concatenate([[0], cuts, [Nz]])
    layer_eps = [float(rng.choice(eps_choices)) for _ in range(n_layers)]
    eps_r = np.empty(g.shape_cells)
    for L in range(n_layers):
        eps_r[:, :, bounds[L]:bounds[L + 1]] = layer_eps[L]
    occ = np.zeros(g.shape_cells, dtype=bool)  # cells claimed incl. a 1-cell gap
    boxes = []
    lo, hi = side
    tries = 0
    while len(boxes) < n_boxes:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place the conductor boxes; enlarge the grid")
        s = [int(rng.integers(lo, min(hi, N - 2) + 1)) for N in (Nx, Ny, Nz)]
        o = [int(rng.integers(1, N - sd)) for N, sd in zip((Nx, Ny, Nz), s)]
        sl = tuple(slice(max(a - 1, 0), a + b + 1) for a, b in zip(o, s))
        if occ[sl].any():
            continue
        occ[sl] = True
        boxes.append((o[0], o[0] + s[0], o[1], o[1] + s[1], o[2], o[2] + s[2]))
    nid = g.node_ids()
    terms = []
    for m, (i0, i1, j0, j1, k0, k1) in enumerate(boxes):
        terms.append(Terminal(f"c{m}", nid[i0:i1 + 1, j0:j1 + 1, k0:k1 + 1].ravel()))
    return Extraction(g, eps_r * eps0, terms, boxes, layer_eps, seed)


def c2_geometry(seed: int = 0, nodes=(192, 192, 96)) -> Extraction:
    """192x192x96 nodes, x/y 50 nm, z log-graded 20 -> 400 nm, three layers with
    eps_r from {2.7, 3.9, 7.5}, eight boxes of 4-40 cells per side."""
    return layered_boxes(nodes=nodes, seed=seed)


def multi_conductor(nodes=(128, 128, 128), seed: int = 0, h: float = 20e-9,
                    n_terms=(2, 8), eps_range=(2.5, 8.0), n_layers=(1, 4),
                    side=None, max_tries=100000) -> Extraction:
    """Seeded multi-conductor extraction on a uniform grid: ``n_layers`` z
    layers with eps_r drawn uniformly from ``eps_range``, and a number of
    rectangular conductor boxes drawn from ``n_terms``.  Boxes have sides of
    ``side`` cells (default 4 .. N/4) and at least one free cell between
    them.  Each box is a terminal owning every node of its closed box."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = nodes
    Nx, Ny, Nz = nx - 1, ny - 1, nz - 1
    g = RectilinearGrid.uniform([(0, Nx * h), (0, Ny * h), (0, Nz * h)], (Nx, Ny, Nz))
    nl = int(rng.integers(n_layers[0], n_layers[1] + 1))
    cuts = np.sort(rng.choice(np.arange(1, Nz), size=nl - 1, replace=False)) if nl > 1 else []
    bounds = np.concatenate([[0], cuts, [Nz]]).astype(int)
    layer_eps = [float(rng.uniform(*eps_range)) for _ in range(nl)]
    eps_r = np.empty(g.shape_cells)
    for L in range(nl):
        eps_r[:, :, bounds[L]:bounds[L + 1]] = layer_eps[L]
    nt = int(rng.integers(n_terms[0], n_terms[1] + 1))
    lo, hi = side if side is not None else (4, max(5, min(Nx, Ny, Nz) // 4))
    occ = np.zeros(g.shape_cells, dtype=bool)
    boxes = []
    tries = 0
    while len(boxes) < nt:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place the conductor boxes; enlarge the grid")
        s = [int(rng.integers(lo, min(hi, N - 2) + 1)) for N in (Nx, Ny, Nz)]
        o = [int(rng.integers(1, N - sd)) for N, sd in zip((Nx, Ny, Nz), s)]
        sl = tuple(slice(max(a - 1, 0), a + b + 1) for a, b in zip(o, s))
        if occ[sl].any():
            continue
        occ[sl] = True
        boxes.append((o[0], o[0] + s[0], o[1], o[1] + s[1], o[2], o[2] + s[2]))
    nid = g.node_ids()
    terms = [Terminal(f"c{m}", nid[i0:i1 + 1, j0:j1 + 1, k0:k1 + 1].ravel())
             for m, (i0, i1, j0, j1, k0, k1) in enumerate(boxes)]
    return Extraction(g, eps_r * eps0, terms, boxes, layer_eps, seed)


def batch_instances(count: int = 64, seed: int = 0, n_min: int = 32 ** 3, n_max: int = 160 ** 3,
                    max_aspect: float = 4.0, n_terms=(2, 6)):
    """Seeded extraction problems of uneven size.  Node counts are drawn
    uniformly in N (not in side length) from [n_min, n_max]; the three side
    ratios are drawn log-uniformly in [1/2, 2] around the cube (so the
    largest over the smallest side is at most ``max_aspect`` = 4); each
    instance is a ``multi_conductor`` geometry with 2-6 terminals and layered
    eps_r in [2.5, 8]."""
    rng = np.random.default_rng(seed)
    out = []
    half = np.sqrt(max_aspect)
    for i in range(count):
        n = float(rng.uniform(n_min, n_max))
        a = np.exp(rng.uniform(np.log(1 / half), np.log(half), 3))
        c = (n / np.prod(a)) ** (1.0 / 3.0)
        sides = tuple(int(max(16, round(c * v))) for v in a)
        m = min(sides)
        ex = multi_conductor(sides, seed=int(rng.integers(1 << 30)), n_terms=n_terms,
                             side=(3, max(4, m // 4)))
        out.append(ex)
    return out


def eqs_two_phase(n: int = 256, seed: int = 0, h: float = 10e-9, fill: float = 0.2,
                  f0: float = 1e9, side=None, max_tries: int = 100000):
    """Seeded two-phase EQS problem on a uniform n^3-node grid: copper boxes
    (sigma 5.8e7 S/m, eps_r 3.9) placed at random until they fill ``fill`` of
    the volume, in SiO2 (sigma 0, eps_r 3.9).  Two boxes with a free gap are
    the terminals: "drv" at sin(2 pi f0 t) V and "gnd" at 0 V; the other
    copper floats.  Returns (grid, eps_cell, sigma_cell, [drv, gnd])."""
    rng = np.random.default_rng(seed)
    N = n - 1
    g = RectilinearGrid.uniform([(0, N * h)] * 3, (N, N, N))
    cu = np.zeros(g.shape_cells, dtype=bool)
    lo, hi = side if side is not None else (max(3, N // 24), max(4, N // 8))
    boxes = []
    tries = 0
    while cu.mean() < fill:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not reach the copper fill")
        s = [int(rng.integers(lo, hi + 1)) for _ in range(3)]
        o = [int(rng.integers(1, N - sd)) for sd in s]
        cu[o[0]:o[0] + s[0], o[1]:o[1] + s[1], o[2]:o[2] + s[2]] = True
        boxes.append((o, s))
    # terminals: two boxes whose closed node sets (plus a one-cell gap) are disjoint
    nid = g.node_ids()
    def nodes_of(b):
        (i, j, k), (a, bb, c) = b
        return nid[i:i + a + 1, j:j + bb + 1, k:k + c + 1].ravel()
    drv_box = boxes[0]
    gnd_box = None
    for b in boxes[1:]:
        (i, j, k), (a, bb, c) = b
        (I, J, K), (A, B, C) = drv_box
        if (i > I + A + 1 or I > i + a + 1 or j > J + B + 1 or J > j + bb + 1 or
                k > K + C + 1 or K > k + c + 1):
            gnd_box = b
            break
    if gnd_box is None:
        raise RuntimeError("no disjoint ground box")
    eps = np.full(g.shape_cells, 3.9 * eps0)
    sigma = np.where(cu, 5.8e7, 0.0)
    drv = Terminal("drv", nodes_of(drv_box), voltage=lambda t: float(np.sin(2 * np.pi * f0 * t)))
    gnd = Terminal("gnd", nodes_of(gnd_box), voltage=0.0)
    return g, eps, sigma, [drv, gnd]
