"""CPU oracle of the learned multilevel preconditioner -- TEST INFRASTRUCTURE ONLY.

The preconditioner has no counterpart in the reference, whose CG is
preconditioned by ILU or Jacobi (fieldspice/solvers/poisson.py:466-484), so
this module restates the GPU library's own definitions from their mathematics.

* ``conv3d`` and its data and weight gradients: a 3x3x3 zero-padded
  convolution written tap by tap as ``y += shift(x) @ W[tap]``.
* ``net_forward`` / ``net_backward``: the 11-layer network (16/32/64
  channels, mean pooling, nearest upsampling, additive skips before the ReLU).
  ``bf16=True`` rounds to BF16 where the device stores BF16: inputs, weights,
  every stored activation and gradient.  Sums are float64.
* ``Hierarchy``: Galerkin coarse operators ``P^T A P`` as sparse matrix
  products of box-aggregation prolongations.
* ``Hierarchy.cycle``: the smoothing/restriction recursion of the cycle;
  ``loss_grad``: the energy-norm loss of the whole cycle and its gradient by
  reverse mode; ``bottom_loss_grad``: the energy functional of the bottom
  system, which the network minimises at the exact bottom solve.

Pinning (tests/test_nn_oracle.py): the backward passes against central finite
differences in float64; the Galerkin operators against the dense triple
product; the recursion of the cycle against its closed form
``M = omega D^-1 + S omega D^-1 + S P M_c P^T S^T`` (``S = I - omega D^-1 A``),
including symmetry for a linear bottom; the bottom loss against finite
differences and its minimiser against a direct solve of the bottom system.

Only ``tests/`` import this module.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

CIN = (16, 16, 16, 32, 32, 64, 64, 32, 32, 16, 16)
COUT = (16, 16, 32, 32, 64, 64, 32, 32, 16, 16, 16)


# --------------------------------------------------------------------------
# BF16
# --------------------------------------------------------------------------
def bf16(x) -> np.ndarray:
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _r(x, on):
    return bf16(x) if on else np.asarray(x, dtype=np.float64)


# --------------------------------------------------------------------------
# parameters: flat FP32 vector, per layer W[27][cin][cout] then b[cout]
# --------------------------------------------------------------------------
def split_params(p):
    p = np.asarray(p, dtype=np.float64)
    out, off = [], 0
    for ci, co in zip(CIN, COUT):
        W = p[off:off + 27 * ci * co].reshape(27, ci, co)
        off += 27 * ci * co
        b = p[off:off + co]
        off += co
        out.append((W, b))
    assert off == p.size, (off, p.size)
    return out


def num_params():
    return sum(27 * ci * co + co for ci, co in zip(CIN, COUT))


def join_grads(gs):
    return np.concatenate([np.concatenate([gW.ravel(), gb.ravel()]) for gW, gb in gs])


# --------------------------------------------------------------------------
# convolution on channels-last (nb, nz, ny, nx, C) tensors; tap =
# (dz+1)*9 + (dy+1)*3 + (dx+1)
# --------------------------------------------------------------------------
def _taps():
    for t in range(27):
        yield t, t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1


def shift(x, dx, dy, dz):
    """s[..., k, j, i, :] = x[..., k+dz, j+dy, i+dx, :] (zero outside)."""
    nb, nz, ny, nx, c = x.shape
    s = np.zeros_like(x)
    zs = slice(max(0, -dz), nz - max(0, dz))
    ys = slice(max(0, -dy), ny - max(0, dy))
    xs = slice(max(0, -dx), nx - max(0, dx))
    zt = slice(max(0, dz), nz - max(0, -dz))
    yt = slice(max(0, dy), ny - max(0, -dy))
    xt = slice(max(0, dx), nx - max(0, -dx))
    s[:, zs, ys, xs, :] = x[:, zt, yt, xt, :]
    return s


def conv3d(x, W, b=None):
    y = np.zeros(x.shape[:-1] + (W.shape[2],))
    for t, dx, dy, dz in _taps():
        y += shift(x, dx, dy, dz) @ W[t]
    if b is not None:
        y = y + b
    return y


def conv3d_data_grad(h, W):
    """d/dx of sum(h * conv3d(x, W))."""
    g = np.zeros(h.shape[:-1] + (W.shape[1],))
    for t, dx, dy, dz in _taps():
        g += shift(h, -dx, -dy, -dz) @ W[t].T
    return g


def conv3d_weight_grad(x, h):
    cin, cout = x.shape[-1], h.shape[-1]
    gW = np.zeros((27, cin, cout))
    hf = h.reshape(-1, cout)
    for t, dx, dy, dz in _taps():
        gW[t] = shift(x, dx, dy, dz).reshape(-1, cin).T @ hf
    return gW


def _child_counts(shape_f):
    nb, nz, ny, nx, _ = shape_f
    cnt = lambda n: np.array([min(2 * I + 1, n - 1) - 2 * I + 1 for I in range((n + 1) // 2)])
    return (cnt(nz)[:, None, None] * cnt(ny)[None, :, None] * cnt(nx)[None, None, :])


def pool(x):
    nb, nz, ny, nx, c = x.shape
    Z, Y, X = (nz + 1) // 2, (ny + 1) // 2, (nx + 1) // 2
    s = np.zeros((nb, Z, Y, X, c))
    for a in range(2):
        for bb in range(2):
            for cc in range(2):
                part = x[:, a::2, bb::2, cc::2, :]
                s[:, :part.shape[1], :part.shape[2], :part.shape[3], :] += part
    return s / _child_counts(x.shape)[None, :, :, :, None]


def pool_bwd(gc, shape_f):
    cnt = _child_counts(shape_f)[None, :, :, :, None]
    return up(gc / cnt, shape_f)


def up(xc, shape_f):
    nb, nz, ny, nx, _ = shape_f
    return xc[:, np.arange(nz) // 2][:, :, np.arange(ny) // 2][:, :, :, np.arange(nx) // 2]


def up_bwd(gf, shape_c):
    s = np.zeros(shape_c)
    for a in range(2):
        for bb in range(2):
            for cc in range(2):
                part = gf[:, a::2, bb::2, cc::2, :]
                s[:, :part.shape[1], :part.shape[2], :part.shape[3], :] += part
    return s


# --------------------------------------------------------------------------
# network
# --------------------------------------------------------------------------
def net_forward(params, X, bf=True):
    """Activations dict of the forward pass; ``acts['Y']`` is the output."""
    L = split_params(params)
    Wr = [(_r(W, bf), b) for W, b in L]
    relu = lambda v: np.maximum(v, 0.0)
    A = {"X0": _r(X, bf)}
    c = lambda i, x, skip=None, act=True: _r(
        (relu if act else (lambda v: v))(conv3d(x, Wr[i][0], Wr[i][1]) +
                                         (0.0 if skip is None else skip)), bf)
    A["A1"] = c(0, A["X0"])
    A["A2"] = c(1, A["A1"])
    A["P2"] = _r(pool(A["A2"]), bf)
    A["A3"] = c(2, A["P2"])
    A["A4"] = c(3, A["A3"])
    A["P4"] = _r(pool(A["A4"]), bf)
    A["A5"] = c(4, A["P4"])
    A["A6"] = c(5, A["A5"])
    A["U6"] = up(A["A6"], A["A4"].shape[:-1] + (64,))
    A["A7"] = c(6, A["U6"], A["A4"])
    A["A8"] = c(7, A["A7"])
    A["U8"] = up(A["A8"], A["A2"].shape[:-1] + (32,))
    A["A9"] = c(8, A["U8"], A["A2"])
    A["A10"] = c(9, A["A9"])
    A["Y"] = c(10, A["A10"], act=False)
    return A


def net_backward(params, A, dY, bf=True):
    """Gradient of sum(dY * Y) with respect to the flat parameters."""
    L = split_params(params)
    Wr = [_r(W, bf) for W, _ in L]
    m = lambda t, ref: np.where(ref > 0, t, 0.0)
    g = [None] * 11

    def wg(i, x, h):
        g[i] = (conv3d_weight_grad(x, h), h.reshape(-1, h.shape[-1]).sum(axis=0))

    dY = _r(dY, bf)
    wg(10, A["A10"], dY)
    H10 = _r(m(conv3d_data_grad(dY, Wr[10]), A["A10"]), bf)
    wg(9, A["A9"], H10)
    H9 = _r(m(conv3d_data_grad(H10, Wr[9]), A["A9"]), bf)
    wg(8, A["U8"], H9)
    dU8 = _r(conv3d_data_grad(H9, Wr[8]), bf)
    H8 = _r(m(up_bwd(dU8, A["A8"].shape), A["A8"]), bf)
    wg(7, A["A7"], H8)
    H7 = _r(m(conv3d_data_grad(H8, Wr[7]), A["A7"]), bf)
    wg(6, A["U6"], H7)
    dU6 = _r(conv3d_data_grad(H7, Wr[6]), bf)
    H6 = _r(m(up_bwd(dU6, A["A6"].shape), A["A6"]), bf)
    wg(5, A["A5"], H6)
    H5 = _r(m(conv3d_data_grad(H6, Wr[5]), A["A5"]), bf)
    wg(4, A["P4"], H5)
    dP4 = _r(conv3d_data_grad(H5, Wr[4]), bf)
    H4 = _r(m(pool_bwd(dP4, A["A4"].shape) + H7, A["A4"]), bf)
    wg(3, A["A3"], H4)
    H3 = _r(m(conv3d_data_grad(H4, Wr[3]), A["A3"]), bf)
    wg(2, A["P2"], H3)
    dP2 = _r(conv3d_data_grad(H3, Wr[2]), bf)
    H2 = _r(m(pool_bwd(dP2, A["A2"].shape) + H9, A["A2"]), bf)
    wg(1, A["A1"], H2)
    H1 = _r(m(conv3d_data_grad(H2, Wr[1]), A["A1"]), bf)
    wg(0, A["X0"], H1)
    return join_grads(g)


def init_params(seed=0):
    """He-normal weights (output layer x0.1), zero biases (any fixed set of
    parameters serves the tests; this one only gives sensible magnitudes)."""
    rng = np.random.default_rng(seed)
    out = []
    for i, (ci, co) in enumerate(zip(CIN, COUT)):
        sd = np.sqrt(2.0 / (27 * ci)) * (0.1 if i == 10 else 1.0)
        out.append(np.concatenate([rng.normal(0, sd, 27 * ci * co), rng.normal(0, 0.05, co)]))
    return np.concatenate(out).astype(np.float32)


# --------------------------------------------------------------------------
# aggregation hierarchy and cycle (vectors in C order of (nx, ny, nz))
# --------------------------------------------------------------------------
def prolongation(shape, groups, active):
    """Box-aggregation P (fine C-order index -> coarse C-order index), rows
    of inactive fine nodes empty."""
    gx, gy, gz = (np.asarray(g) for g in groups)
    cshape = (int(gx[-1]) + 1, int(gy[-1]) + 1, int(gz[-1]) + 1)
    I, J, K = np.meshgrid(gx, gy, gz, indexing="ij")
    col = ((I * cshape[1] + J) * cshape[2] + K).ravel()
    n = int(np.prod(shape))
    m = np.asarray(active).ravel()
    rows = np.arange(n)
    P = sp.csr_matrix((np.ones(int(m.sum())), (rows[m], col[m])), shape=(n, int(np.prod(cshape))))
    return P, cshape


class Hierarchy:
    """Levels of the cycle for an equilibrated level-0 matrix ``A0`` (CSR,
    identity rows on inactive nodes) on a node grid ``shape`` (C order)."""

    def __init__(self, A0, active0, shape, groups, omega=0.8):
        self.omega = float(omega)
        self.A = [sp.csr_matrix(A0)]
        self.act = [np.asarray(active0, bool).ravel()]
        self.shape = [tuple(shape)]
        self.P = []
        for gs in groups:
            P, cshape = prolongation(self.shape[-1], gs, self.act[-1])
            Ac = (P.T @ self.A[-1] @ P).tocsr()
            act = np.asarray(P.sum(axis=0)).ravel() > 0
            Ac = (Ac + sp.diags((~act).astype(float))).tocsr()
            self.P.append(P)
            self.A.append(Ac)
            self.act.append(act)
            self.shape.append(cshape)
        self.dinv = [np.where(a, 1.0 / A.diagonal(), 0.0) for A, a in zip(self.A, self.act)]
        self.K = len(groups)

    # bottom features in C order: couplings normalised by sqrt(d_n d_m) along
    # +x, +y, +z, and activity; dh = 1/sqrt(d) on active nodes
    def bottom_features(self):
        A, act, shape = self.A[-1], self.act[-1], self.shape[-1]
        nx, ny, nz = shape
        d = A.diagonal()
        dh = np.where(act, 1.0 / np.sqrt(np.where(act, d, 1.0)), 0.0)
        idx = np.arange(nx * ny * nz).reshape(shape)
        feat = np.zeros((nx * ny * nz, 4))
        Ad = A.todok() if A.shape[0] < 4000 else None
        for a, (di, dj, dk) in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
            src = idx[:nx - di, :ny - dj, :nz - dk].ravel()
            dst = idx[di:, dj:, dk:].ravel()
            c = -np.asarray(A[src, dst]).ravel()
            feat[src, a] = np.where(act[src] & act[dst], c * dh[src] * dh[dst], 0.0)
        feat[:, 3] = act.astype(float)
        return feat, dh

    def to_grid(self, v):
        """C-order bottom vector(s) -> (..., nz, ny, nx)."""
        nx, ny, nz = self.shape[-1]
        v = np.asarray(v)
        return np.swapaxes(v.reshape(v.shape[:-1] + (nx, ny, nz)), -1, -3)

    def from_grid(self, g):
        nx, ny, nz = self.shape[-1]
        return np.swapaxes(g, -1, -3).reshape(g.shape[:-3] + (nx * ny * nz,))

    def bottom_net(self, params, rK, bf=True):
        """e_K for a batch rK (nb, nK) through the network; also returns what
        the backward pass needs."""
        feat, dh = self.bottom_features()
        nb = rK.shape[0]
        s = np.sqrt(np.mean((rK * dh) ** 2, axis=1))
        inv = np.where(s > 0, 1.0 / np.where(s > 0, s, 1.0), 0.0)
        X = np.zeros(rK.shape + (16,))
        X[..., 0] = rK * dh * inv[:, None]
        X[..., 1:5] = feat[None]
        Xg = np.stack([np.stack([self.to_grid(X[b, :, c]) for c in range(16)], axis=-1)
                       for b in range(nb)])
        A = net_forward(params, Xg, bf)
        Y0 = np.stack([self.from_grid(A["Y"][b, ..., 0]) for b in range(nb)])
        e = s[:, None] * dh[None] * Y0
        return e, (A, s, dh)

    def cycle(self, r0, bottom):
        """M(r0) for a batch (nb, n0); ``bottom(rK) -> eK``."""
        om = self.omega
        rs = [np.atleast_2d(r0)]
        for l in range(self.K):
            z1 = om * self.dinv[l] * rs[l]
            rs.append((self.P[l].T @ (rs[l] - (self.A[l] @ z1.T).T).T).T)
        e = bottom(rs[-1])
        for l in range(self.K - 1, -1, -1):
            z2 = om * self.dinv[l] * rs[l] + (self.P[l] @ e.T).T
            e = z2 + om * self.dinv[l] * (rs[l] - (self.A[l] @ z2.T).T)
        return e

    def restrict_chain(self, g0):
        """P_{K-1}^T S_{K-1}^T ... P_0^T S_0^T g0: the adjoint of the up-sweep."""
        om = self.omega
        g = np.atleast_2d(g0)
        for l in range(self.K):
            t = g - om * (self.A[l] @ (self.dinv[l] * g).T).T
            g = (self.P[l].T @ t.T).T
        return g

    def closed_form(self, Mc):
        """Dense matrix of the cycle for a dense bottom matrix Mc (levels
        combined from the bottom up)."""
        M = np.asarray(Mc, dtype=np.float64)
        om = self.omega
        for l in range(self.K - 1, -1, -1):
            A = self.A[l].toarray()
            Di = np.diag(self.dinv[l])
            S = np.eye(A.shape[0]) - om * Di @ A
            P = self.P[l].toarray()
            M = om * Di + S @ (om * Di) + S @ P @ M @ P.T @ S.T
        return M

    def loss_grad(self, params, e0, bf=True):
        """Per-element loss ||e - M(A e)||_A^2/||e||_A^2 and the gradient of
        its batch mean with respect to the parameters."""
        e0 = np.atleast_2d(e0)
        A0 = self.A[0]
        r0 = (A0 @ e0.T).T
        E = np.einsum("bi,bi->b", e0, r0)
        store = {}

        def bottom(rK):
            e, aux = self.bottom_net(params, rK, bf)
            store["aux"] = aux
            return e
        z = self.cycle(r0, bottom)
        Az = (A0 @ z.T).T
        loss = (np.einsum("bi,bi->b", z, Az) - 2 * np.einsum("bi,bi->b", z, r0)) / E + 1.0
        nb = e0.shape[0]
        g0 = 2.0 * (Az - r0) / (E[:, None] * nb)
        gK = self.restrict_chain(g0)
        A, s, dh = store["aux"]
        dY = np.zeros(A["Y"].shape)
        for b in range(nb):
            dY[b, ..., 0] = self.to_grid(s[b] * dh * gK[b])
        grad = net_backward(params, A, dY, bf)
        return loss, grad, z

    def bottom_loss_grad(self, params, e0, bf=True):
        """Bottom loss (y^T A_K y - 2 y^T r_K) / (r_K^T D_K^-1 r_K) per element,
        r_K the down-sweep of r = A e, y the network's bottom error; and the
        gradient of its batch mean."""
        e0 = np.atleast_2d(e0)
        r0 = (self.A[0] @ e0.T).T
        rK = self.restrict_chain(r0)
        y, (A, s, dh) = self.bottom_net(params, rK, bf)
        AK = self.A[-1]
        w = (AK @ y.T).T
        c = np.einsum("bi,bi->b", rK * self.dinv[-1], rK)
        loss = (np.einsum("bi,bi->b", y, w) - 2 * np.einsum("bi,bi->b", y, rK)) / c
        nb = e0.shape[0]
        g = 2.0 * (w - rK) / (c[:, None] * nb)
        dY = np.zeros(A["Y"].shape)
        for b in range(nb):
            dY[b, ..., 0] = self.to_grid(s[b] * dh * g[b])
        return loss, net_backward(params, A, dY, bf), y
