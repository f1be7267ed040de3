"""The production kernels against the oracle: FP32 Krylov iterations run the
stencil with W nodes per lane (`k_stencil_dotw`, W the widest of 4, 3, 2 that
divides the rows), FP64 ones the 2-nodes-per-lane stencil and residual on
even rows, and the learned flexible CG its 4-wide level-0 restriction and
prolongation on rows a multiple of 4.  Each solve reports the kernel width it ran, and the test
asserts it, so a fallback to the one-node-per-lane kernels cannot pass
unnoticed.  C and phi are compared with the oracle's direct solve
(fieldspice/solvers/poisson.py:1185-1249 restated) at 1e-9.
"""

import numpy as np
import pytest

import fs_oracle as O
from fieldspice_amd import workloads
from fieldspice_amd.solvers.base import SolverConfig
from fieldspice_amd.solvers.poisson import PoissonSolver

pytestmark = pytest.mark.gpu

# (label, geometry): x extents a multiple of 4 (4-wide stencil and cycle), of
# 3 only (3 nodes per lane, one-node cycle) and of 2 only (2 nodes per lane)
CASES = {
    "layered 48x40x28, 5 boxes": lambda: workloads.layered_boxes(
        nodes=(48, 40, 28), n_boxes=5, side=(2, 6), seed=21),
    "layered 264x16x12, two 256-node x tiles": lambda: workloads.layered_boxes(
        nodes=(264, 16, 12), n_boxes=4, side=(2, 5), seed=21),
    "uniform 64x32x24": lambda: workloads.multi_conductor(
        (64, 32, 24), seed=23, n_terms=(3, 5), side=(3, 6)),
    "layered 198x20x14, 3 nodes per lane, two x tiles": lambda: workloads.layered_boxes(
        nodes=(198, 20, 14), n_boxes=4, side=(2, 5), seed=22),
    "layered 50x28x22, 2 nodes per lane": lambda: workloads.layered_boxes(
        nodes=(50, 28, 22), n_boxes=4, side=(2, 5), seed=23),
}


def widths(solver, nx):
    """(stencil width, level-0 cycle width) the library selects by default."""
    if solver == "hip_pcg":
        return (2 if nx % 2 == 0 else 1, 0)
    w = 4 if nx % 4 == 0 else 3 if nx % 3 == 0 else 2 if nx % 2 == 0 else 1
    return (w, (4 if nx % 4 == 0 else 1) if solver == "hip_fcg" else 0)

_oracle_cache = {}


def _oracle(label):
    if label not in _oracle_cache:
        ex = CASES[label]()
        Cr, _, phis = O.capacitance_matrix(ex.grid, ex.eps_cell, [t.nodes for t in ex.terminals],
                                           method="direct", return_phi=True)
        _oracle_cache[label] = (ex, Cr, phis)
    return _oracle_cache[label]


@pytest.mark.parametrize("label", list(CASES))
@pytest.mark.parametrize("solver", ["hip_pcg", "hip_mixed", "hip_fcg"])
def test_wide_kernels_capacitance_matches_oracle(gpu_lib, label, solver):
    print(f"[gpu case] wide kernels: {label}, {solver}", flush=True)
    ex, Cr, phis = _oracle(label)
    # tol 1e-12: the potentials' error against the direct solve is bounded by
    # the condition number times the relative residual, so a 1e-10 residual
    # alone leaves a 1e-9 comparison of phi without margin
    ps = PoissonSolver(ex.grid, ex.eps_cell, SolverConfig(linear_solver=solver, tol=1e-12))
    rep = {}
    C, phi = ps.capacitance_matrix(ex.terminals, report=rep, return_phi=True)
    lin = rep["linear"]
    assert (lin["stencil_width"], lin["cycle_width"]) == widths(solver, ex.grid.shape_nodes[0]), lin
    assert lin["residual_eq"] <= 1e-12, lin
    err = np.abs(C - Cr).max() / np.abs(Cr).max()
    assert err <= 1e-9, err
    for j in range(len(ex.terminals)):
        assert np.abs(phi[j] - phis[j]).max() <= 1e-9, j
    assert np.all(np.diag(C) > 0) and np.all(C - np.diag(np.diag(C)) <= 0)
