"""fs_sys_solve_noalloc (the workspace in the library's static scratch)
against fs_sys_solve on the same systems: two calls in a row, the second on
a smaller system with fewer right-hand sides, in each precision; the
system's own workspace still works afterwards; oversize requests fail
cleanly before any kernel runs."""

import numpy as np
import pytest

from fieldspice_amd import _lib, workloads
from fieldspice_amd.precond import LearnedPreconditioner

pytestmark = pytest.mark.gpu


def _problem(nodes, seed, nrhs):
    ex = workloads.layered_boxes(nodes=nodes, n_boxes=3, side=(2, 4), seed=seed)
    g = ex.grid
    L = _lib.DeviceOperator.from_cells(g, ex.eps_cell)
    mask = np.zeros(g.n_nodes, np.uint8)
    for t in ex.terminals:
        mask[t.nodes] = 1
    sysd = _lib.DeviceSystem(L, mask)
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((nrhs, g.n_nodes)) * 1e-17
    fixed = mask.astype(bool)
    B[:, fixed] = rng.uniform(0.0, 1.0, (nrhs, int(fixed.sum())))
    return ex, L, sysd, B


def _agree(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("precision", [_lib.FS_PREC_FP64, _lib.FS_PREC_MIXED, _lib.FS_PREC_FCG])
def test_noalloc_solve_matches_fs_sys_solve(gpu_lib, precision):
    print(f"[gpu case] no-allocation solve, precision {precision}", flush=True)
    big = _problem((48, 40, 28), 31, 5)
    small = _problem((20, 18, 14), 32, 2)
    pcs = []
    if precision == _lib.FS_PREC_FCG:
        lp = LearnedPreconditioner()
        pcs = [lp.attach(p[2], p[0].grid) for p in (big, small)]
    prm = _lib.DeviceSystem.params(1e-10, precision=precision)
    for (ex, L, sysd, B), use_x0 in ((big, True), (small, False)):
        x0 = 0.5 * B if use_x0 else None
        xr, str_, rcr = sysd.solve(B, prm, x0=x0)
        xn, stn, rcn = sysd.solve(B, prm, x0=x0, noalloc=True)
        assert rcr == 0 and rcn == 0, (rcr, rcn)
        err = _agree(xn, xr)
        print(f"  {ex.grid.shape_nodes}, {B.shape[0]} RHS: iterations {stn['iterations']} / "
              f"{str_['iterations']}, max rel difference {err:.2e}, bitwise "
              f"{np.array_equal(xn, xr)}", flush=True)
        assert err <= 1e-12
        assert stn["iterations"] == str_["iterations"]
        assert stn["residual_eq"] <= 1e-10
        # the system's own workspace is back in place
        xa, sta, _ = sysd.solve(B, prm, x0=x0)
        assert _agree(xa, xr) <= 1e-12
    del pcs


def test_noalloc_rejects_requests_beyond_its_scratch(gpu_lib):
    print("[gpu case] no-allocation solve: oversize request", flush=True)
    g_nodes = (96, 96, 80)     # 737,280 nodes: 8 RHS need ~400 MB of scratch
    ex, L, sysd, B = _problem(g_nodes, 33, 8)
    with pytest.raises(_lib.LibraryError, match="scratch"):
        sysd.solve(B, _lib.DeviceSystem.params(1e-10), noalloc=True)
    x, st, rc = sysd.solve(B[:1], _lib.DeviceSystem.params(1e-10), noalloc=True)
    assert rc == 0 and st["residual_eq"] <= 1e-10
