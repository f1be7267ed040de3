"""Pinning of the learned-preconditioner oracle (oracle/fs_nn_oracle.py) on the
CPU: backward passes against central finite differences, the Galerkin coarse
operator against its closed form on a uniform Laplacian, the cycle recursion
against its closed matrix form, and the adjoint used by training."""

import numpy as np
import pytest
import scipy.sparse as sp

import fs_nn_oracle as N
import fs_oracle as O
from fieldspice_amd.grid import RectilinearGrid
from fieldspice_amd.precond import axis_groups, hierarchy_groups


def _equilibrated(grid, eps_cell, fixed):
    L = O.nodal_laplacian(grid, O.cell_to_edge(grid, eps_cell)).tocsr()
    free = ~fixed
    d = L.diagonal()
    dh = np.where(free, 1.0 / np.sqrt(d), 0.0)
    Dm = sp.diags(free.astype(float))
    A = sp.diags(dh) @ (Dm @ L @ Dm) @ sp.diags(dh) + sp.diags(fixed.astype(float))
    return A.tocsr(), free


def _small_system(shape=(7, 6, 5), seed=0):
    rng = np.random.default_rng(seed)
    g = RectilinearGrid.uniform([(0, 1), (0, 1.2), (0, 0.8)], tuple(s - 1 for s in shape))
    eps = rng.uniform(1.0, 5.0, g.shape_cells)
    fixed = np.zeros(g.n_nodes, bool)
    fixed[:shape[1] * shape[2]] = True            # x = 0 plane
    A, free = _equilibrated(g, eps, fixed)
    return g, A, free


def test_bf16_rounding_ties_to_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0, 1.0 + 2 ** -9])
    # 1+2^-8 is the midpoint between 1 and 1+2^-7 -> even (1); 1+3*2^-8 -> 1+2^-6
    np.testing.assert_array_equal(N.bf16(x), [1.0, 1.0, 1.0 + 2 ** -6, -2.5, 0.0, 1.0])


def test_net_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = N.init_params(3).astype(np.float64)
    X = rng.normal(size=(2, 3, 4, 5, 16))
    dY = rng.normal(size=X.shape)
    A = N.net_forward(p, X, bf=False)
    g = N.net_backward(p, A, dY, bf=False)
    f = lambda q: float(np.sum(dY * N.net_forward(q, X, bf=False)["Y"]))
    for trial in range(3):
        v = rng.normal(size=p.size)
        h = 1e-6
        fd = (f(p + h * v) - f(p - h * v)) / (2 * h)
        assert abs(fd - g @ v) <= 1e-6 * max(1.0, abs(fd)), (trial, fd, g @ v)


def test_conv_gradients_match_finite_differences():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(1, 3, 4, 5, 16))
    W = rng.normal(size=(27, 16, 32))
    h = rng.normal(size=(1, 3, 4, 5, 32))
    f = lambda xx, WW: float(np.sum(h * N.conv3d(xx, WW)))
    gx, gW = N.conv3d_data_grad(h, W), N.conv3d_weight_grad(x, h)
    vx, vW = rng.normal(size=x.shape), rng.normal(size=W.shape)
    e = 1e-6
    assert abs((f(x + e * vx, W) - f(x - e * vx, W)) / (2 * e) - np.sum(gx * vx)) < 1e-5
    assert abs((f(x, W + e * vW) - f(x, W - e * vW)) / (2 * e) - np.sum(gW * vW)) < 1e-5


def test_conv_is_a_plain_correlation():
    # one-hot kernel at tap (dx,dy,dz) = (+1,0,-1): y[k,j,i] = x[k-1, j, i+1]
    rng = np.random.default_rng(3)
    x = rng.normal(size=(1, 4, 3, 5, 16))
    W = np.zeros((27, 16, 16))
    t = (-1 + 1) * 9 + (0 + 1) * 3 + (1 + 1)
    W[t] = np.eye(16)
    y = N.conv3d(x, W)
    np.testing.assert_array_equal(y[:, 1:, :, :-1], x[:, :-1, :, 1:])
    assert not y[:, 0].any() and not y[:, :, :, -1].any()


def test_pool_up_adjoints():
    rng = np.random.default_rng(4)
    xf = rng.normal(size=(2, 5, 3, 4, 16))
    xc = rng.normal(size=(2, 3, 2, 2, 16))
    assert np.isclose(np.sum(N.pool(xf) * xc), np.sum(xf * N.pool_bwd(xc, xf.shape)))
    assert np.isclose(np.sum(N.up(xc, xf.shape) * xf), np.sum(xc * N.up_bwd(xf, xc.shape)))
    # pool of a constant is that constant (partial boxes average fewer nodes)
    np.testing.assert_allclose(N.pool(np.ones_like(xf)), 1.0)


def test_axis_groups_pairs_uniform_and_merge_fine_spacing():
    hd = np.r_[0.5, np.ones(8), 0.5]
    np.testing.assert_array_equal(axis_groups(hd, 2.0), [0, 0, 1, 1, 2, 2, 3, 3, 4, 4])
    # widths 0.1 merge up to 1.0; the wide ones stay single
    g = axis_groups([0.1] * 10 + [2.0, 2.0], 1.0)
    np.testing.assert_array_equal(g, [0] * 10 + [1, 2])
    lv = hierarchy_groups((np.ones(64), np.ones(64), np.ones(64)), max_bottom=8 ** 3)
    assert [int(gs[0][-1]) + 1 for gs in lv] == [32, 16, 8]


def test_galerkin_operator_of_a_uniform_laplacian():
    # unit-diagonal 7-point operator with couplings 1/6 (periodic-free
    # interior): 2x2x2 boxes give diag 8 - 2*12/6 = 4 and couplings 4/6
    n = 8
    idx = np.arange(n ** 3).reshape(n, n, n)
    rows, cols = [], []
    for a in range(3):
        s = [slice(None)] * 3
        t = [slice(None)] * 3
        s[a], t[a] = slice(0, n - 1), slice(1, n)
        rows.append(idx[tuple(s)].ravel())
        cols.append(idx[tuple(t)].ravel())
    r, c = np.concatenate(rows), np.concatenate(cols)
    C = sp.csr_matrix((np.full(r.size, 1 / 6), (r, c)), shape=(n ** 3, n ** 3))
    A = sp.eye(n ** 3) - C - C.T
    pairs = np.arange(n) // 2
    H = N.Hierarchy(A, np.ones(n ** 3, bool), (n, n, n), [(pairs, pairs, pairs)])
    Ac = H.A[1].toarray()
    m = 4 // 2 * (16 + 4 + 1)     # an interior coarse node (1,1,1) of 4^3
    assert np.isclose(Ac[m, m], 4.0)
    for off in (1, 4, 16):
        assert np.isclose(Ac[m, m + off], -4 / 6) and np.isclose(Ac[m, m - off], -4 / 6)


@pytest.mark.parametrize("omega", [0.7, 0.9])
def test_cycle_recursion_matches_closed_form(omega):
    g, A, free = _small_system()
    groups = hierarchy_groups((g.ax.hd, g.ay.hd, g.az.hd), max_bottom=20)
    assert len(groups) >= 2
    H = N.Hierarchy(A, free, g.shape_nodes, groups, omega)
    Mc = np.diag(omega * H.dinv[-1])
    M = H.closed_form(Mc)
    rng = np.random.default_rng(5)
    R = rng.normal(size=(3, A.shape[0])) * free
    Z = H.cycle(R, lambda rK: omega * H.dinv[-1] * rK)
    np.testing.assert_allclose(Z, R @ M.T, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(M, M.T, atol=1e-12)      # symmetric for a linear bottom
    assert np.all(np.linalg.eigvalsh(M[np.ix_(free, free)]) > 0)


def test_restriction_chain_is_the_adjoint_of_the_up_sweep():
    g, A, free = _small_system(seed=1)
    groups = hierarchy_groups((g.ax.hd, g.ay.hd, g.az.hd), max_bottom=20)
    H = N.Hierarchy(A, free, g.shape_nodes, groups)
    rng = np.random.default_rng(6)
    eK = rng.normal(size=(1, A.shape[0] if False else H.A[-1].shape[0])) * H.act[-1]
    gvec = rng.normal(size=(1, A.shape[0])) * free
    up = H.cycle(np.zeros((1, A.shape[0])), lambda rK: eK)   # r = 0: only the up-sweep
    assert np.isclose(float(np.sum(up * gvec)), float(np.sum(eK * H.restrict_chain(gvec))))


def test_loss_gradient_matches_finite_differences():
    g, A, free = _small_system((9, 8, 6), seed=2)
    groups = hierarchy_groups((g.ax.hd, g.ay.hd, g.az.hd), max_bottom=60)
    H = N.Hierarchy(A, free, g.shape_nodes, groups)
    rng = np.random.default_rng(7)
    p = N.init_params(1).astype(np.float64)
    e = rng.normal(size=(2, A.shape[0])) * free
    loss, grad, _ = H.loss_grad(p, e, bf=False)
    assert np.all(np.isfinite(loss))
    f = lambda q: float(np.mean(H.loss_grad(q, e, bf=False)[0]))
    for trial in range(2):
        v = rng.normal(size=p.size)
        h = 1e-6
        fd = (f(p + h * v) - f(p - h * v)) / (2 * h)
        assert abs(fd - grad @ v) <= 1e-5 * max(1.0, abs(fd)), (fd, grad @ v)


def test_bottom_loss_gradient_and_minimiser():
    g, A, free = _small_system((9, 8, 6), seed=3)
    groups = hierarchy_groups((g.ax.hd, g.ay.hd, g.az.hd), max_bottom=60)
    H = N.Hierarchy(A, free, g.shape_nodes, groups)
    rng = np.random.default_rng(8)
    p = N.init_params(2).astype(np.float64)
    e = rng.normal(size=(2, A.shape[0])) * free
    loss, grad, y = H.bottom_loss_grad(p, e, bf=False)
    f = lambda q: float(np.mean(H.bottom_loss_grad(q, e, bf=False)[0]))
    v = rng.normal(size=p.size)
    h = 1e-6
    fd = (f(p + h * v) - f(p - h * v)) / (2 * h)
    assert abs(fd - grad @ v) <= 1e-5 * max(1.0, abs(fd)), (fd, grad @ v)
    # the functional is bounded below by its value at the exact bottom solve
    rK = H.restrict_chain((A @ e.T).T)
    AK = H.A[-1].toarray()
    yK = np.linalg.solve(AK, rK.T).T
    c = np.einsum("bi,bi->b", rK * H.dinv[-1], rK)
    best = (np.einsum("bi,bi->b", yK, (AK @ yK.T).T) - 2 * np.einsum("bi,bi->b", yK, rK)) / c
    assert np.all(loss >= best - 1e-12) and np.all(best < 0)
