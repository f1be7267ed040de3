"""Profiling of the Krylov loop (fs_sys_set_profiling): every level-0 kernel
kind of a learned flexible-CG solve is timed by HIP events and its
algorithmic bytes counted on the device, results are bit-identical to the
unprofiled solve, and the counted bytes follow the per-kind formulas of the
library from the launch counts.  Eight right-hand sides split the level-0
prolongation into two launches (start event on the first, stop event on the
second); five take it in one."""

import numpy as np
import pytest

from fieldspice_amd import _lib, workloads
from fieldspice_amd.solvers.base import SolverConfig, check_terminals
from fieldspice_amd.solvers.linear import GridLinearSystem

pytestmark = pytest.mark.gpu


def _cap(ex, method):
    g = ex.grid
    L = _lib.DeviceOperator.from_cells(g, ex.eps_cell)
    terms = check_terminals(ex.terminals, g)
    tid = np.full(g.n_nodes, -1, dtype=np.int32)
    for k, t in enumerate(terms):
        tid[t.nodes] = k
    ls = GridLinearSystem(L, tid >= 0, SolverConfig(linear_solver=method, tol=1e-10), grid=g)
    return L, ls, ls.sys.prepare_capacitance(tid, len(terms))


@pytest.mark.parametrize("nt", [5, 8])
def test_profiled_fcg_times_every_level0_kernel(gpu_lib, nt):
    print(f"[gpu case] profiled learned FCG, {nt} terminals", flush=True)
    ex = workloads.multi_conductor((64, 48, 40), seed=31, n_terms=(nt, nt), side=(3, 6))
    assert len(ex.terminals) == nt
    L, ls, cap = _cap(ex, "hip_fcg")
    params = ls.params()
    C0, _, st0, rc0 = cap.solve(L, params)
    ls.sys.set_profiling(True)
    C1, _, st1, rc1 = cap.solve(L, params)
    ls.sys.set_profiling(False)
    assert rc0 == 0 and rc1 == 0
    assert np.array_equal(C0, C1)  # the events change nothing the solve computes
    assert st1["iterations"] == st0["iterations"] and st1["converged"] == 1
    assert all(t == 0 for t in st0["kernel_time"])  # unprofiled: nothing recorded
    names = _lib.PROF_NAMES
    n = ex.grid.n_nodes
    for k, name in enumerate(names):
        assert st1["kernel_launches"][k] > 0, name
        assert st1["kernel_time"][k] > 0, name
        per = st1["kernel_bytes"][k] / st1["kernel_launches"][k]
        assert per > 0, name
    # every iteration runs the stencil, the two vector updates and one cycle;
    # each cycle start adds one restriction and one prolongation
    it = st1["kernel_launches"][names.index("stencil")]
    assert st1["kernel_launches"][names.index("direction")] == it
    assert st1["kernel_launches"][names.index("residual")] == it
    assert st1["kernel_launches"][names.index("restrict")] == \
        st1["kernel_launches"][names.index("prolong")] >= it
    assert st1["stencil_launches"] == it and st1["stencil_time"] == st1["kernel_time"][0]
    # FP32 stencil bytes: 12 n per launch with work + 8 n per live RHS, at most all nt live
    b = st1["kernel_bytes"][names.index("stencil")]
    assert b <= it * n * (12 + 8 * nt) * (1 + 1e-12)
    assert b >= it * n * (12 + 8 * 1) * (1 - 1e-12)
    # direction update per live RHS and launch: p and x read and written, z
    # read (FP16 over a 4-wide cycle: 18 n bytes, else 20 n)
    assert st1["cycle_width"] == 4  # 64-node rows, x and y in pairs
    bd = st1["kernel_bytes"][names.index("direction")]
    assert bd <= it * n * 18 * nt * (1 + 1e-12) and bd % (18 * n) == 0


def test_stream_probe_times_a_mix(gpu_lib):
    print("[gpu case] streaming probe", flush=True)
    n = 1 << 22
    t = _lib.stream_probe(n, 8, 5, reps=5)
    rate = n * 4.0 * 13 / t / 1e9
    assert 100.0 < rate < 20000.0, rate  # GB/s: finite and physical
    with pytest.raises(_lib.LibraryError):
        _lib.stream_probe(n, 7, 5)  # not a supported mix
    with pytest.raises(_lib.LibraryError):
        _lib.stream_probe(n + 2, 8, 5)  # not a multiple of 4
