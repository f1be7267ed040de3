"""Learned multilevel preconditioner for flexible CG (``linear_solver="hip_fcg"``).

The reference preconditions its CG with ILU or Jacobi
(``fieldspice/solvers/poisson.py:466-484``).  This package adds a multilevel
preconditioner that lives on the GPU:

* an aggregation hierarchy of the prepared, equilibrated system.  Each level
  groups consecutive nodes along each axis; a box of the grouping is one node
  of the next level.  With piecewise-constant prolongation the Galerkin
  product ``P^T A P`` is again a 7-point operator, so every level streams
  like the fine one.
* groupings chosen from the node spacings.  An axis is coarsened only where
  its spacing is small next to the finest of the others.  On the
  log-graded meshes of layered extraction problems this keeps the coarse
  cells near-isotropic, and without it point smoothing stalls on the
  anisotropy.
* damped Jacobi smoothing on every level.
* at the bottom, a 3-level convolutional network (``_lib.DeviceNet``,
  BF16 MFMA kernels).  It maps the bottom residual and the bottom operator's
  normalised couplings to an approximate bottom error.

The network is trained on the GPU from residual/error pairs of FP64 solves
(``train`` below).  The loss is the energy norm of the error left by the
whole cycle: ``||e - M(r)||_A^2 / ||e||_A^2``.
"""

from __future__ import annotations

import os
import warnings

import numpy as np

from . import _lib

__all__ = ["axis_groups", "hierarchy_groups", "LearnedPreconditioner", "default_weights_path",
           "load_net", "save_net", "uniform_widths", "DEFAULT_OMEGA", "DEFAULT_MAX_BOTTOM"]

# Jacobi weight of the cycle's smoothing.  Below 1 the smoother damps every
# mode of the equilibrated operator (whose spectrum lies in (0, 2)), so the
# cycle stays symmetric positive definite.  0.95 against 0.8 (the weight the
# shipped network was trained with): 119 vs 128 flexible-CG iterations at
# 256^3, 84 vs 86 on the graded 192x192x96 case, 89 vs 97 at 160^3.
DEFAULT_OMEGA = 0.95
DEFAULT_MAX_BOTTOM = 8_000
WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")


def default_weights_path() -> str:
    return os.path.join(WEIGHTS_DIR, "precond_net.npz")


def axis_groups(widths, target: float) -> np.ndarray:
    """Greedy grouping of consecutive nodes (dual widths ``widths``) into
    groups whose summed width stays within ``target``.  Every group holds at
    least one node.  Returns the group id of each node (0, then steps of 0 or
    1)."""
    w = np.asarray(widths, dtype=np.float64).ravel()
    gid = np.empty(w.size, dtype=np.int32)
    cur, acc = 0, 0.0
    lim = target * (1.0 + 1e-9)
    for i, wi in enumerate(w):
        if acc > 0.0 and acc + wi > lim:
            cur += 1
            acc = 0.0
        gid[i] = cur
        acc += wi
    return gid


def hierarchy_groups(widths, max_bottom: int = DEFAULT_MAX_BOTTOM, max_levels: int = 8,
                     min_levels: int = 1):
    """Per-level groupings ``[(gx, gy, gz), ...]`` for node dual widths
    ``widths = (wx, wy, wz)``.

    At each level the target group width is twice the finest median width of
    the axes that still have more than one node.  Coarsening stops once the
    bottom has at most ``max_bottom`` nodes (and at least ``min_levels``
    levels exist), or when no axis coarsens any more.
    """
    ws = [np.asarray(w, dtype=np.float64).ravel() for w in widths]
    levels = []
    for _ in range(max_levels):
        n = int(np.prod([w.size for w in ws]))
        if len(levels) >= min_levels and n <= max_bottom:
            break
        live = [np.median(w) for w in ws if w.size > 1]
        if not live:
            break
        target = 2.0 * min(live)
        gs = [axis_groups(w, target) if w.size > 1 else np.zeros(1, np.int32) for w in ws]
        if all(int(g[-1]) + 1 == w.size for g, w in zip(gs, ws)):
            break
        levels.append(tuple(gs))
        ws = [np.bincount(g, weights=w) for g, w in zip(gs, ws)]
    return levels


def level_shapes(n_nodes_xyz, groups):
    shapes = [tuple(int(v) for v in n_nodes_xyz)]
    for gs in groups:
        shapes.append(tuple(int(g[-1]) + 1 for g in gs))
    return shapes


def save_net(net: "_lib.DeviceNet", path: str, **meta) -> None:
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez(path, params=net.get_params(), num_params=np.int64(net.num_params),
             **{k: np.asarray(v) for k, v in meta.items()})


def load_net(path: str | None = None) -> "_lib.DeviceNet":
    path = path or default_weights_path()
    with np.load(path, allow_pickle=False) as f:
        p = np.asarray(f["params"], dtype=np.float32)
    net = _lib.DeviceNet(seed=0)
    net.set_params(p)
    return net


class LearnedPreconditioner:
    """Builds the device preconditioner of a prepared system.

    ``net``: a ``DeviceNet``.  ``None`` loads the trained weights shipped with
    the package.  ``False`` builds the cycle with one Jacobi step at the
    bottom instead of the network.
    """

    _shared = None

    def __init__(self, net=None, omega: float = DEFAULT_OMEGA,
                 max_bottom: int = DEFAULT_MAX_BOTTOM):
        if net is None:
            if LearnedPreconditioner._shared is None:
                path = default_weights_path()
                if not os.path.exists(path):
                    raise FileNotFoundError(
                        f"{path} is missing: train the preconditioner with "
                        "`python -m fieldspice_amd.train_precond`")
                LearnedPreconditioner._shared = load_net(path)
            net = LearnedPreconditioner._shared
        self.net = None if net is False else net
        self.omega = float(omega)
        self.max_bottom = int(max_bottom)

    def groups_for(self, grid):
        return hierarchy_groups((grid.ax.hd, grid.ay.hd, grid.az.hd), self.max_bottom)

    def build_widths(self, system: "_lib.DeviceSystem", widths) -> "_lib.DevicePreconditioner":
        groups = hierarchy_groups(widths, self.max_bottom)
        if not groups:
            raise ValueError("grid too small for a multilevel preconditioner")
        return _lib.DevicePreconditioner(system, self.net, groups, self.omega)

    def build(self, system: "_lib.DeviceSystem", grid) -> "_lib.DevicePreconditioner":
        return self.build_widths(system, (grid.ax.hd, grid.ay.hd, grid.az.hd))

    def attach_widths(self, system, widths) -> "_lib.DevicePreconditioner":
        pc = self.build_widths(system, widths)
        system.set_precond(pc)
        return pc

    def attach(self, system: "_lib.DeviceSystem", grid) -> "_lib.DevicePreconditioner":
        return self.attach_widths(system, (grid.ax.hd, grid.ay.hd, grid.az.hd))


def uniform_widths(shape_cells):
    """Dual widths of a unit-spaced grid with ``shape_cells`` cells."""
    out = []
    for nc in shape_cells:
        w = np.ones(int(nc) + 1)
        w[0] = w[-1] = 0.5
        out.append(w)
    return tuple(out)


def warn_untrained():
    warnings.warn("no trained preconditioner weights: the cycle uses one Jacobi step at "
                  "its bottom", RuntimeWarning, stacklevel=3)
