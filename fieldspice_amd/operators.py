"""Host-side material averaging and edge-metric helpers.

The per-edge operator weights of the solve are built on the GPU
(``fs_op_create_cells``).  These helpers cover the node-centred averages the
thermal step needs and the edge-property -> weight conversion of
``edge_mass`` (``fieldspice/operators.py:270-315``).
"""

from __future__ import annotations

import numpy as np

__all__ = ["cell_to_node", "edge_weights_from_property", "node_volume_vector"]


def node_volume_vector(grid) -> np.ndarray:
    """Dual (box) volume per node, flat C order."""
    return grid.node_volumes().ravel()


def cell_to_node(grid, prop) -> np.ndarray:
    """Volume-weighted average of a per-cell property onto nodes
    (``fieldspice/operators.py:270-296``): each node averages the <= 8 cells
    around it, weighted by the octant of each cell inside its dual box."""
    prop = np.asarray(prop, dtype=float)
    if prop.shape != grid.shape_cells:
        raise ValueError(f"prop must have shape {grid.shape_cells}")
    padded = np.pad(prop, 1, mode="edge")
    half = [(np.concatenate([[0.0], 0.5 * h]), np.concatenate([0.5 * h, [0.0]]))
            for h in (grid.hx, grid.hy, grid.hz)]
    nx, ny, nz = grid.shape_nodes
    num = np.zeros(grid.shape_nodes)
    den = np.zeros(grid.shape_nodes)
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                w = np.multiply.outer(np.multiply.outer(half[0][a], half[1][b]), half[2][c])
                num += padded[a:a + nx, b:b + ny, c:c + nz] * w
                den += w
    return num / np.where(den > 0, den, 1.0)


def edge_weights_from_property(grid, prop_edge) -> np.ndarray:
    """``prop * A_dual / L`` per edge, concatenated [x, y, z]."""
    prop_edge = np.asarray(prop_edge, dtype=float).ravel()
    L = np.concatenate([a.ravel() for a in grid.edge_lengths()])
    A = np.concatenate([a.ravel() for a in grid.edge_dual_areas()])
    if prop_edge.size != L.size:
        raise ValueError(f"expected {L.size} edge values, got {prop_edge.size}")
    return prop_edge * A / L
