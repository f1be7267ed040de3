"""Batched extraction of uneven systems in one captured graph, and the stencil
tile autotuner, against per-system solves on the GPU."""

import numpy as np
import pytest

from fieldspice_amd import _lib
from fieldspice_amd.precond import LearnedPreconditioner, default_weights_path
from fieldspice_amd.workloads import multi_conductor

pytestmark = pytest.mark.gpu

SHAPES = [(24, 20, 18), (30, 16, 22), (20, 20, 20), (36, 24, 16), (17, 33, 19)]


def _instances(with_pc):
    out = []
    for k, shp in enumerate(SHAPES):
        ex = multi_conductor(shp, seed=500 + k, n_terms=(2, 4), side=(3, 5))
        g = ex.grid
        L = _lib.DeviceOperator.from_cells(g, ex.eps_cell)
        tid = np.full(g.n_nodes, -1, np.int32)
        for m, t in enumerate(ex.terminals):
            tid[t.nodes] = m
        sysd = _lib.DeviceSystem(L, (tid >= 0).astype(np.uint8))
        kz = sysd.autotune(len(ex.terminals), reps=3)
        assert 1 <= kz <= g.shape_nodes[2]
        pc = LearnedPreconditioner().attach(sysd, g) if with_pc else None
        cap = sysd.prepare_capacitance(tid, len(ex.terminals))
        out.append((ex, L, sysd, pc, cap))
    return out


def test_batch_matches_per_system_solves():
    print("[gpu case] batched extraction vs per-system FP64", flush=True)
    assert default_weights_path()
    inst = _instances(with_pc=True)
    ref = [cap.solve(L, _lib.DeviceSystem.params(1e-11))[0] for ex, L, sysd, pc, cap in inst]
    for prec, chunk in ((_lib.FS_PREC_FCG, 4), (_lib.FS_PREC_FP64, 16), (_lib.FS_PREC_MIXED, 8)):
        batch = _lib.DeviceBatch(nstreams=3)
        for ex, L, sysd, pc, cap in inst:
            batch.add(cap, L)
        Cs, t, its, st, rc = batch.solve(_lib.DeviceSystem.params(1e-10, precision=prec), chunk)
        assert rc == 0 and st["converged"] == 1, (prec, st)
        assert np.all(t > 0) and np.all(its > 0)
        for C, R in zip(Cs, ref):
            assert np.abs(C - R).max() <= 1e-8 * np.abs(R).max(), (prec, np.abs(C - R).max())
        print(f"precision {prec}: iterations {its.tolist()}, finish times "
              f"{np.round(t * 1e3, 1).tolist()} ms", flush=True)
        batch.close()
    # the systems are usable one by one after a batch (own scalars and stream)
    ex, L, sysd, pc, cap = inst[0]
    C0 = cap.solve(L, _lib.DeviceSystem.params(1e-10, precision=_lib.FS_PREC_FCG))[0]
    assert np.abs(C0 - ref[0]).max() <= 1e-8 * np.abs(ref[0]).max()


def test_batch_rejects_fcg_without_preconditioner():
    print("[gpu case] batch argument checks", flush=True)
    inst = _instances(with_pc=False)[:2]
    batch = _lib.DeviceBatch(nstreams=2)
    for ex, L, sysd, pc, cap in inst:
        batch.add(cap, L)
    with pytest.raises(_lib.LibraryError):
        batch.solve(_lib.DeviceSystem.params(1e-10, precision=_lib.FS_PREC_FCG))
    with pytest.raises(_lib.LibraryError):
        batch.add(inst[0][4], inst[0][1])        # same system twice
