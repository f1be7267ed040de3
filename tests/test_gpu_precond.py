"""Learned multilevel preconditioner on the GPU against the CPU oracle
(oracle/fs_nn_oracle.py): the MFMA network forward and backward, the cycle
with and without the network, the training loss and gradient; then training,
flexible CG and its FP64 fallback."""

import numpy as np
import pytest
import scipy.sparse as sp

import fs_nn_oracle as N
import fs_oracle as O
from fieldspice_amd import _lib
from fieldspice_amd.grid import RectilinearGrid
from fieldspice_amd.precond import hierarchy_groups

pytestmark = pytest.mark.gpu


def _case(shape=(20, 14, 11), seed=0, plane=True, box=((0, 1), (0, 0.8), (0, 0.6))):
    rng = np.random.default_rng(seed)
    g = RectilinearGrid.uniform(list(box), tuple(s - 1 for s in shape))
    eps = rng.uniform(1.0, 6.0, g.shape_cells)
    fixed = np.zeros(g.n_nodes, bool)
    if plane:
        fixed[:shape[1] * shape[2]] = True
    nid = g.node_ids()
    fixed[nid[shape[0] // 2:shape[0] // 2 + 3, 3:6, 2:5].ravel()] = True
    L = O.nodal_laplacian(g, O.cell_to_edge(g, eps)).tocsr()
    free = ~fixed
    d = L.diagonal()
    dh = np.where(free, 1.0 / np.sqrt(d), 0.0)
    A = (sp.diags(dh) @ (sp.diags(free * 1.0) @ L @ sp.diags(free * 1.0)) @ sp.diags(dh)
         + sp.diags(fixed * 1.0)).tocsr()
    w = O.edge_weights(g, O.cell_to_edge(g, eps))
    op = _lib.DeviceOperator.from_edges(g, w)
    sysd = _lib.DeviceSystem(op, fixed.astype(np.uint8))
    return g, A, free, np.sqrt(d) * free, op, sysd, L, fixed


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def test_net_forward_matches_oracle():
    print("[gpu case] net forward vs oracle", flush=True)
    rng = np.random.default_rng(0)
    p = N.init_params(5)
    net = _lib.DeviceNet(seed=1)
    net.set_params(p)
    X = rng.normal(size=(2, 5, 6, 21, 16)).astype(np.float32)
    Y = net.eval(X)
    ref = N.net_forward(p, X, bf=True)["Y"]
    assert _rel(Y, ref) < 5e-3, _rel(Y, ref)
    assert np.mean(Y == ref) > 0.9          # mostly bit-equal BF16 values


def test_net_backward_matches_oracle():
    print("[gpu case] net backward vs oracle", flush=True)
    rng = np.random.default_rng(1)
    p = N.init_params(6)
    net = _lib.DeviceNet(seed=2)
    net.set_params(p)
    X = rng.normal(size=(2, 4, 9, 18, 16)).astype(np.float32)
    dY = rng.normal(size=X.shape).astype(np.float32)
    Y, grad = net.eval(X, dY)
    A = N.net_forward(p, X, bf=True)
    ref = N.net_backward(p, A, dY, bf=True)
    assert _rel(grad, ref) < 1e-2, _rel(grad, ref)
    # every layer's weight and bias gradient individually
    off = 0
    for ci, co in zip(N.CIN, N.COUT):
        nw = 27 * ci * co
        assert _rel(grad[off:off + nw], ref[off:off + nw]) < 3e-2
        assert _rel(grad[off + nw:off + nw + co], ref[off + nw:off + nw + co]) < 3e-2
        off += nw + co


@pytest.mark.parametrize("shape,omega", [((20, 14, 11), 0.8), ((69, 13, 9), 0.7)])
def test_cycle_with_jacobi_bottom_matches_oracle(shape, omega):
    print(f"[gpu case] cycle (Jacobi bottom) vs oracle {shape}", flush=True)
    g, A, free, dh, op, sysd, L, fixed = _case(shape)
    groups = hierarchy_groups((g.ax.hd, g.ay.hd, g.az.hd), max_bottom=150)
    pc = _lib.DevicePreconditioner(sysd, None, groups, omega)
    H = N.Hierarchy(A, free, g.shape_nodes, groups, omega)
    for lv in range(len(groups) + 1):
        assert pc.level_dims(lv) == H.shape[lv]
    rng = np.random.default_rng(2)
    R = rng.normal(size=(3, g.n_nodes)) * free
    Z = pc.apply(R)
    ref = H.cycle(R, lambda rK: omega * H.dinv[-1] * rK)
    assert _rel(Z, ref) < 2e-6, _rel(Z, ref)
    assert np.all(Z[:, fixed] == 0.0)


@pytest.mark.parametrize("nx,nrhs", [(256, 5), (64, 7)])
def test_cycle_coarse_levels_on_tiles_match_oracle(nx, nrhs):
    # equal spacings: level 0 and level 1 (nx/2 x 6 x 5) are grouped in pairs,
    # so both run the 4-wide z-march kernels (level 1 with its diagonal, and
    # a 32-node row at nx = 64); 5 and 7 right-hand sides use the exact-count
    # instantiations and the two half-width prolongation launches
    print(f"[gpu case] cycle on 4-wide tiles at levels 0 and 1 vs oracle, nx {nx}", flush=True)
    box = ((0, 1), (0, 11 / (nx - 1)), (0, 9 / (nx - 1)))
    g, A, free, dh, op, sysd, L, fixed = _case((nx, 12, 10), seed=4, box=box)
    groups = hierarchy_groups((g.ax.hd, g.ay.hd, g.az.hd), max_bottom=150)
    pc = _lib.DevicePreconditioner(sysd, None, groups, 0.8)
    H = N.Hierarchy(A, free, g.shape_nodes, groups, 0.8)
    assert pc.level_dims(1) == (nx // 2, 6, 5)
    assert pc.nlevels >= 2
    rng = np.random.default_rng(5)
    R = rng.normal(size=(nrhs, g.n_nodes)) * free
    Z = pc.apply(R)
    ref = H.cycle(R, lambda rK: 0.8 * H.dinv[-1] * rK)
    assert _rel(Z, ref) < 2e-6, _rel(Z, ref)
    assert np.all(Z[:, fixed] == 0.0)


def test_cycle_with_network_matches_oracle():
    print("[gpu case] cycle (network bottom) vs oracle", flush=True)
    g, A, free, dh, op, sysd, L, fixed = _case(seed=3)
    groups = hierarchy_groups((g.ax.hd, g.ay.hd, g.az.hd), max_bottom=200)
    p = N.init_params(7)
    net = _lib.DeviceNet(seed=0)
    net.set_params(p)
    pc = _lib.DevicePreconditioner(sysd, net, groups, 0.8)
    H = N.Hierarchy(A, free, g.shape_nodes, groups, 0.8)
    rng = np.random.default_rng(4)
    R = rng.normal(size=(2, g.n_nodes)) * free
    Z = pc.apply(R)
    ref = H.cycle(R, lambda rK: H.bottom_net(p, rK, bf=True)[0])
    assert _rel(Z, ref) < 1e-3, _rel(Z, ref)


def test_loss_and_gradient_match_oracle():
    print("[gpu case] training loss/gradient vs oracle", flush=True)
    g, A, free, dh, op, sysd, L, fixed = _case((18, 12, 10), seed=5)
    groups = hierarchy_groups((g.ax.hd, g.ay.hd, g.az.hd), max_bottom=150)
    p = N.init_params(8)
    net = _lib.DeviceNet(seed=0)
    net.set_params(p)
    pc = _lib.DevicePreconditioner(sysd, net, groups, 0.8)
    H = N.Hierarchy(A, free, g.shape_nodes, groups, 0.8)
    rng = np.random.default_rng(6)
    e = rng.normal(size=(3, g.n_nodes)) * free          # node-space errors
    loss, grad = pc.loss_grad(e, kind=_lib.FS_LOSS_CYCLE)
    lref, gref, _ = H.loss_grad(p, e * dh, bf=True)
    np.testing.assert_allclose(loss, lref, rtol=2e-3)
    assert _rel(grad, gref) < 2e-2, _rel(grad, gref)
    loss, grad = pc.loss_grad(e, kind=_lib.FS_LOSS_BOTTOM)
    lref, gref, _ = H.bottom_loss_grad(p, e * dh, bf=True)
    np.testing.assert_allclose(loss, lref, rtol=2e-3)
    assert _rel(grad, gref) < 2e-2, _rel(grad, gref)


def _lifted(L, fixed, xD):
    """Right-hand side after Dirichlet elimination (what LinearSystem.solve
    takes): -L x_D on free rows, the Dirichlet values on fixed rows."""
    b = -(L @ xD)
    b[fixed] = xD[fixed]
    return b


def _pairs_from_solver(sysd, L, fixed, terminal_nodes, ks=(2, 5, 10, 20, 40)):
    """Residual/error pairs from the FP64 solver: e_k = x* - x_k of CG runs
    stopped after k iterations, for unit excitations of each terminal."""
    out = []
    for nodes in terminal_nodes:
        xD = np.zeros(L.shape[0])
        xD[nodes] = 1.0
        b = _lifted(L, fixed, xD)
        P = _lib.DeviceSystem.params(1e-13)
        xs, st, rc = sysd.solve(b, P)
        for k in ks:
            xk, st, rc = sysd.solve(b, _lib.DeviceSystem.params(1e-13, maxiter=k))
            out.append(xs - xk)
    return np.array(out)


def test_training_lowers_loss_and_fcg_beats_the_untrained_cycle():
    print("[gpu case] training + flexible CG", flush=True)
    shape = (40, 40, 40)
    rng = np.random.default_rng(11)
    g = RectilinearGrid.uniform([(0, 1), (0, 1), (0, 1)], tuple(s - 1 for s in shape))
    eps = np.where(rng.uniform(size=g.shape_cells) < 0.5, 2.5, 7.5)
    nid = g.node_ids()
    terms = [nid[5:12, 5:12, 5:12].ravel(), nid[25:34, 20:30, 22:35].ravel(),
             nid[8:30, 30:34, 3:8].ravel()]
    fixed = np.zeros(g.n_nodes, bool)
    for t in terms:
        fixed[t] = True
    w = O.edge_weights(g, O.cell_to_edge(g, eps))
    L = O.nodal_laplacian(g, O.cell_to_edge(g, eps)).tocsr()
    op = _lib.DeviceOperator.from_edges(g, w)
    sysd = _lib.DeviceSystem(op, fixed.astype(np.uint8))
    groups = hierarchy_groups((g.ax.hd, g.ay.hd, g.az.hd), max_bottom=3000)
    net = _lib.DeviceNet(seed=3)
    pc = _lib.DevicePreconditioner(sysd, net, groups, 0.8)
    pc.add_pairs(_pairs_from_solver(sysd, L, fixed, terms))
    losses = pc.train(300, batch=4, lr=2e-3, seed=1)
    print(f"loss first 30 {losses[:30].mean():.4f}  last 30 {losses[-30:].mean():.4f}", flush=True)
    assert np.all(np.isfinite(losses))
    # the bottom energy is negative and decreases towards the exact solve's
    assert losses[-30:].mean() < losses[:30].mean() - 0.1 * abs(losses[:30].mean())
    # flexible CG with the trained network, against the cycle with one Jacobi
    # step at the bottom and plain FP64 CG, on lifted right-hand sides
    xD = np.zeros((2, g.n_nodes))
    xD[0, terms[1]] = 1.0
    xD[1, terms[2]] = 1.0
    b = np.stack([_lifted(L, fixed, v) for v in xD])
    x64, st64, _ = sysd.solve(b, _lib.DeviceSystem.params(1e-10))
    sysd.set_precond(pc)
    x_l, st_l, rc = sysd.solve(b, _lib.DeviceSystem.params(1e-10, precision=_lib.FS_PREC_FCG))
    assert rc == 0 and st_l["fallbacks"] == 0
    base = _lib.DevicePreconditioner(sysd, None, groups, 0.8)
    sysd.set_precond(base)
    x_j, st_j, rc = sysd.solve(b, _lib.DeviceSystem.params(1e-10, precision=_lib.FS_PREC_FCG))
    assert rc == 0
    for x in (x_l, x_j):
        assert np.abs(x - x64).max() <= 1e-8 * np.abs(x64).max()
    print(f"iterations: FP64 CG {st64['iterations']}, FCG+Jacobi bottom {st_j['iterations']}, "
          f"FCG+network {st_l['iterations']}", flush=True)
    assert st_l["iterations"] < st_j["iterations"] < st64["iterations"]


def test_fcg_falls_back_to_fp64_cg():
    print("[gpu case] flexible CG fallback", flush=True)
    g, A, free, dh, op, sysd, L, fixed = _case(seed=9)
    groups = hierarchy_groups((g.ax.hd, g.ay.hd, g.az.hd), max_bottom=150)
    pc = _lib.DevicePreconditioner(sysd, None, groups, 0.8)
    sysd.set_precond(pc)
    xD = np.zeros(g.n_nodes)
    xD[fixed] = 1.0
    xD[:g.shape_nodes[1] * g.shape_nodes[2]] = 0.0      # the x = 0 plane grounded
    b = _lifted(L, fixed, xD)
    x_ref, st_ref, _ = sysd.solve(b, _lib.DeviceSystem.params(1e-11))
    assert st_ref["iterations"] > 2
    x, st, rc = sysd.solve(b, _lib.DeviceSystem.params(1e-11, maxiter=2,
                                                        precision=_lib.FS_PREC_FCG))
    assert rc == 0 and st["fallbacks"] == 1 and st["converged"] == 1
    assert np.abs(x - x_ref).max() <= 1e-9
    with pytest.raises(_lib.LibraryError):
        sysd.set_precond(None)
        sysd.solve(b, _lib.DeviceSystem.params(1e-11, precision=_lib.FS_PREC_FCG))
