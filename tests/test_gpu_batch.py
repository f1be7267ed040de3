"""Batched extraction of uneven systems in captured graphs under device
control, and the stencil tile autotuner: every batched capacitance matrix
against the oracle's direct
solve (fieldspice/solvers/poisson.py:1185-1249 restated) in all three
precisions, and the full-size 64-system workload through the properties
every Maxwell capacitance matrix has."""

import numpy as np
import pytest

import fs_oracle as O
from fieldspice_amd import _lib, workloads
from fieldspice_amd.precond import LearnedPreconditioner, default_weights_path
from fieldspice_amd.workloads import multi_conductor

pytestmark = pytest.mark.gpu

# x extents mix multiples of 4 (4-wide FP32 kernels) with odd and even others
SHAPES = [(24, 20, 18), (30, 16, 22), (20, 20, 20), (36, 24, 16), (17, 33, 19)]
TOL = 1e-10


def _system(ex, with_pc, lp=None):
    g = ex.grid
    L = _lib.DeviceOperator.from_cells(g, ex.eps_cell)
    tid = np.full(g.n_nodes, -1, np.int32)
    for m, t in enumerate(ex.terminals):
        tid[t.nodes] = m
    sysd = _lib.DeviceSystem(L, (tid >= 0).astype(np.uint8))
    kz = sysd.autotune(len(ex.terminals), reps=3)
    assert 1 <= kz <= g.shape_nodes[2]
    pc = (lp or LearnedPreconditioner()).attach(sysd, g) if with_pc else None
    cap = sysd.prepare_capacitance(tid, len(ex.terminals))
    return L, sysd, pc, cap


def _instances(with_pc):
    out = []
    for k, shp in enumerate(SHAPES):
        ex = multi_conductor(shp, seed=500 + k, n_terms=(2, 4), side=(3, 5))
        out.append((ex,) + _system(ex, with_pc))
    return out


def _sym(C):
    return 0.5 * (C + C.T)


def test_batch_matches_oracle_in_every_precision():
    print("[gpu case] batched extraction vs the oracle, FCG / FP64 / mixed", flush=True)
    assert default_weights_path()
    inst = _instances(with_pc=True)
    ref = [O.capacitance_matrix(ex.grid, ex.eps_cell, [t.nodes for t in ex.terminals],
                                method="direct")[0] for ex, *_ in inst]
    # groups (one stream and host thread each): fewer than the systems, one,
    # and more than the systems (then one system per group)
    for prec, chunk, ns in ((_lib.FS_PREC_FCG, 4, 3), (_lib.FS_PREC_FP64, 16, 1),
                            (_lib.FS_PREC_MIXED, 8, 8)):
        print(f"[gpu case] batch precision {prec}, {ns} groups", flush=True)
        batch = _lib.DeviceBatch(nstreams=ns)
        for ex, L, sysd, pc, cap in inst:
            batch.add(cap, L)
        Cs, t, its, st, rc = batch.solve(_lib.DeviceSystem.params(TOL, precision=prec), chunk)
        assert rc == 0 and st["converged"] == 1, (prec, st)
        # the largest FP64-recomputed residual of every system's every RHS
        assert st["residual_eq"] <= TOL, (prec, st)
        assert np.all(t > 0) and np.all(its > 0)
        # finish times come from the device clock, inside the call
        assert np.all(t <= st["solve_time"]), (t, st["solve_time"])
        if prec != _lib.FS_PREC_FP64:
            # refreshes and cycle starts ran under device control
            assert st["refreshes"] >= 1, st
        if prec == _lib.FS_PREC_FCG:
            # per FP32 cycle one preconditioner application at its start and
            # one after each iteration but the last, so at least one per
            # iteration of the system's slowest right-hand side
            assert st["pc_applies"] >= its.sum(), (st["pc_applies"], its)
        for k, (C, R) in enumerate(zip(Cs, ref)):
            err = np.abs(_sym(C) - R).max() / np.abs(R).max()
            assert err <= 1e-9, (prec, k, err)
        print(f"precision {prec}: iterations {its.tolist()}, finish times "
              f"{np.round(t * 1e3, 1).tolist()} ms, replacements {st['replacements']}",
              flush=True)
        batch.close()
    # the systems are usable one by one after a batch (own scalars and stream)
    ex, L, sysd, pc, cap = inst[0]
    C0, _, st0, _ = cap.solve(L, _lib.DeviceSystem.params(TOL, precision=_lib.FS_PREC_FCG))
    assert np.abs(_sym(C0) - ref[0]).max() <= 1e-9 * np.abs(ref[0]).max()
    assert st0["residual_eq"] <= TOL


def test_batch_redoes_failed_fcg_systems_in_fp64():
    """Flexible CG capped at 3 iterations cannot converge: the device control
    finishes every system with the iteration-cap flag, and the batch redoes
    each in plain FP64 CG (default cap), whose C still meets the oracle."""
    print("[gpu case] batch: FCG capped, FP64 redo", flush=True)
    inst = _instances(with_pc=True)[:3]
    ref = [O.capacitance_matrix(ex.grid, ex.eps_cell, [t.nodes for t in ex.terminals],
                                method="direct")[0] for ex, *_ in inst]
    batch = _lib.DeviceBatch(nstreams=2)
    for ex, L, sysd, pc, cap in inst:
        batch.add(cap, L)
    Cs, t, its, st, rc = batch.solve(
        _lib.DeviceSystem.params(TOL, maxiter=3, precision=_lib.FS_PREC_FCG), 4)
    batch.close()
    assert rc == 0 and st["converged"] == 1, st
    assert st["fallbacks"] == len(inst), st
    assert st["residual_eq"] <= TOL, st
    assert np.all(t > 0) and np.all(t <= st["solve_time"])
    for k, (C, R) in enumerate(zip(Cs, ref)):
        err = np.abs(_sym(C) - R).max() / np.abs(R).max()
        assert err <= 1e-9, (k, err)


def test_full_size_batch_properties():
    """64 systems of 32^3 .. 160^3 nodes (154M nodes, 257 right-hand sides):
    symmetric, zero row sums (no grounded wall), Maxwell signs, and every
    right-hand side at its FP64-recomputed residual.  (The oracle's direct
    solve takes minutes per system at these sizes; the shrunk systems of the
    test above carry the oracle comparison.)"""
    print("[gpu case] full-size 64-system batch: setup", flush=True)
    exs = workloads.batch_instances(64, seed=0)
    lp = LearnedPreconditioner()
    inst = []
    for k, ex in enumerate(exs):
        inst.append((ex,) + _system(ex, True, lp))
        if k % 16 == 15:
            print(f"  {k + 1} systems prepared", flush=True)
    batch = _lib.DeviceBatch(nstreams=16)
    for ex, L, sysd, pc, cap in inst:
        batch.add(cap, L)
    print("[gpu case] full-size 64-system batch: learned FCG solve", flush=True)
    Cs, t, its, st, rc = batch.solve(_lib.DeviceSystem.params(TOL, precision=_lib.FS_PREC_FCG),
                                     4)
    batch.close()
    assert rc == 0 and st["converged"] == 1 and st["fallbacks"] == 0, st
    assert st["residual_eq"] <= TOL, st
    assert len(Cs) == 64 and sum(C.shape[0] for C in Cs) == 257
    # A column's sum is minus the charge the solve leaves on the free nodes,
    # sum_n d_n^1/2 r_s,n; at tol 1e-10 over 10^5..10^6 nodes that is ~1e-9
    # of the largest diagonal, so the bounds below are 1e-8.
    worst_asym = worst_rows = 0.0
    for k, C in enumerate(Cs):
        scale = np.abs(np.diag(C)).max()
        worst_asym = max(worst_asym, np.abs(C - C.T).max() / scale)
        S = _sym(C)
        worst_rows = max(worst_rows, np.abs(S.sum(axis=1)).max() / scale)
        assert np.all(np.diag(S) > 0), k
        assert np.all(S - np.diag(np.diag(S)) <= 1e-9 * scale), k
    print(f"max asymmetry {worst_asym:.2e}, max row sum {worst_rows:.2e} of the diagonal, "
          f"iterations max {int(its.max())}, replacements {st['replacements']}", flush=True)
    assert worst_asym <= 1e-8 and worst_rows <= 1e-8


def test_batch_rejects_fcg_without_preconditioner():
    print("[gpu case] batch argument checks", flush=True)
    inst = _instances(with_pc=False)[:2]
    batch = _lib.DeviceBatch(nstreams=2)
    for ex, L, sysd, pc, cap in inst:
        batch.add(cap, L)
    with pytest.raises(_lib.LibraryError):
        batch.solve(_lib.DeviceSystem.params(TOL, precision=_lib.FS_PREC_FCG))
    with pytest.raises(_lib.LibraryError):
        batch.add(inst[0][4], inst[0][1])        # same system twice
