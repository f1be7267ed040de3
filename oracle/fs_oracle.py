"""CPU FP64 oracle for the elliptic field solve -- TEST INFRASTRUCTURE ONLY.

This module restates, in NumPy/SciPy, the algorithm fieldspice runs on the CPU
for ``G^T M G phi = rhs``.  It is the checker that the GPU library is compared
against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg
of ``bench.py`` may import it; the ``fieldspice_amd`` package never does.

Every function cites the reference file:line it follows.  The assembled-matrix
form (CSR ``G^T diag(w) G``, symmetric Dirichlet elimination, Jacobi
equilibration, SciPy CG with ``rtol=tol, atol=0``) is deliberately the
reference's, not the GPU library's matrix-free form, so that agreement between
the two is evidence rather than a restatement of the same code.

How each part is pinned (see the tests named in brackets):

* grid metrics, ``cell_to_edge``, ``edge_weights``: the reference's own
  known-answer tests (single-cell conductance, 15/10/20 area weighting,
  harmonic mode) and invariants (uniform in -> uniform out, dual areas tile
  the cross-section)  [tests/test_oracle.py]
* ``nodal_laplacian``/``apply_dirichlet``: symmetry, constant null vector,
  +-1 incidence, and the closed-form parallel-plate / slab / series-stack
  results to 1e-12..1e-10  [tests/test_oracle.py]
* ``LinearSystem`` CG branch: equal to the direct ``splu`` solve of the same
  system to 1e-9, true residual below tolerance; empty-row and gauge pinning
  reproduce the reference's warnings  [tests/test_oracle.py]
* ``capacitance_matrix``: eps*A/d for the plate (1e-12), series stack, row
  sums, reciprocity; golden fixtures written by the reference itself for a
  graded multi-conductor case  [tests/test_oracle.py, tests/golden/]
* ``eqs_transient``: capacitive divider 2/3, analytic RC step (<2e-4), the
  published BE error 1.216e-3 and first-order ratios, current conservation
  [tests/test_oracle.py]
* ``thermal_*``: parabolic profile, slab R_th, convective R_th, an RC ladder
  built independently from the 1D network  [tests/test_oracle.py]
* ``boltzmann_charge``/``nonlinear_poisson``: the built-in potential of an
  abrupt pn junction, Vt ln(Na Nd/ni^2) = 0.833370 V, to 1e-9 (the
  reference's docs/CONTRACTS.md:455,480-481), a charge-neutral guess, Gauss's
  law over the junction, and a wrong derivative rejected by the line search;
  with merged electrodes, equal to ``poisson_floating`` without space charge
  and the prescribed electrode charge met exactly  [tests/test_oracle.py]
* ``field_region_ac``: the uniform slab's (sigma + j omega eps) A/d, omega C
  of an insulator (the capacitance extraction above), the DC conductance as
  omega -> 0, complex symmetry and zero column sums  [tests/test_oracle.py]
* ``poisson_floating``: a floating plate between driven plates at the
  series-capacitor potential (Q + C1 V)/(C1 + C2), with and without charge;
  the electrode's charge equals the prescribed one; a zero-charge electrode
  in a uniform field leaves the field unchanged  [tests/test_oracle.py]
"""

from __future__ import annotations

import time
import warnings

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

EPS0 = 1.0 / ((4e-7 * np.pi * 1.0000000000552) * 299_792_458.0 ** 2)


# --------------------------------------------------------------------------
# operators (fieldspice/operators.py)
# --------------------------------------------------------------------------
def grad_node_edge(grid) -> sp.csr_matrix:
    """+-1 node->edge incidence, edges ordered [x, y, z] each C-order
    (fieldspice/operators.py:71-91)."""
    nid = np.arange(grid.n_nodes).reshape(grid.shape_nodes)
    heads, tails = [], []
    heads.append(nid[1:, :, :]); tails.append(nid[:-1, :, :])
    heads.append(nid[:, 1:, :]); tails.append(nid[:, :-1, :])
    heads.append(nid[:, :, 1:]); tails.append(nid[:, :, :-1])
    h = np.concatenate([a.ravel() for a in heads])
    t = np.concatenate([a.ravel() for a in tails])
    ne = h.size
    rows = np.concatenate([np.arange(ne), np.arange(ne)])
    cols = np.concatenate([h, t])
    vals = np.concatenate([np.ones(ne), -np.ones(ne)])
    return sp.coo_matrix((vals, (rows, cols)), shape=(ne, grid.n_nodes)).tocsr()


def _halves(h):
    """Half widths of the cells below / above each node plane (0 outside)."""
    return np.concatenate([[0.0], 0.5 * h]), np.concatenate([0.5 * h, [0.0]])


def cell_to_edge(grid, prop, mode: str = "parallel") -> np.ndarray:
    """Quadrant-weighted cell->edge average, concatenated [x, y, z]
    (fieldspice/operators.py:144-237).  An edge is surrounded by up to four
    cells; cell (a, b) of the transverse pair contributes with weight equal to
    the product of its half widths in the two transverse directions.  Missing
    cells (outside the domain) get weight 0."""
    prop = np.asarray(prop, dtype=float)
    if prop.shape != grid.shape_cells:
        raise ValueError(f"prop must have shape {grid.shape_cells}")
    if mode not in ("parallel", "arithmetic", "harmonic", "series", "min", "max"):
        raise ValueError(f"unknown averaging mode {mode!r}")
    p = np.pad(prop, 1, mode="edge")
    hv = [_halves(h) for h in (grid.hx, grid.hy, grid.hz)]
    out = []
    for d in range(3):
        t0, t1 = [a for a in range(3) if a != d]
        shape = grid.shape_edges[d]
        qs, ws = [], []
        for s0 in (0, 1):
            for s1 in (0, 1):
                sl = [slice(None)] * 3
                sl[d] = slice(1, -1)
                sl[t0] = slice(s0, s0 + shape[t0])
                sl[t1] = slice(s1, s1 + shape[t1])
                qs.append(p[tuple(sl)])
                w0 = hv[t0][s0].reshape([-1 if a == t0 else 1 for a in range(3)])
                w1 = hv[t1][s1].reshape([-1 if a == t1 else 1 for a in range(3)])
                ws.append(np.broadcast_to(w0 * w1, shape))
        q = np.stack(qs)
        w = np.stack(ws)
        wsum = w.sum(axis=0)
        safe = np.where(wsum > 0, wsum, 1.0)
        if mode in ("parallel", "arithmetic"):
            e = (q * w).sum(axis=0) / safe
        elif mode in ("harmonic", "series"):
            qq = np.where(q > 0, q, np.finfo(float).tiny)
            e = safe / (w / qq).sum(axis=0)
        elif mode == "min":
            e = q.min(axis=0)
        else:
            e = q.max(axis=0)
        out.append(e.ravel())
    return np.concatenate(out)


def edge_weights(grid, prop_edge) -> np.ndarray:
    """``prop * A_dual / L`` per edge (fieldspice/operators.py:302-315)."""
    prop_edge = np.asarray(prop_edge, dtype=float).ravel()
    L = np.concatenate([a.ravel() for a in grid.edge_lengths()])
    A = np.concatenate([a.ravel() for a in grid.edge_dual_areas()])
    if prop_edge.size != L.size:
        raise ValueError("wrong number of edge values")
    return prop_edge * A / L


def nodal_laplacian(grid, prop_edge, G=None) -> sp.csr_matrix:
    """``G^T diag(prop*A/L) G`` (fieldspice/operators.py:340-351)."""
    G = grad_node_edge(grid) if G is None else G
    return sp.csr_matrix(G.T @ sp.diags(edge_weights(grid, prop_edge)) @ G)


def cell_to_node(grid, prop) -> np.ndarray:
    """Volume-weighted cell->node average (fieldspice/operators.py:270-296)."""
    prop = np.asarray(prop, dtype=float)
    p = np.pad(prop, 1, mode="edge")
    hv = [_halves(h) for h in (grid.hx, grid.hy, grid.hz)]
    nx, ny, nz = grid.shape_nodes
    acc = np.zeros(grid.shape_nodes)
    wacc = np.zeros(grid.shape_nodes)
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                w = (hv[0][a][:, None, None] * hv[1][b][None, :, None]
                     * hv[2][c][None, None, :])
                acc += p[a:a + nx, b:b + ny, c:c + nz] * w
                wacc += w
    return acc / np.where(wacc > 0, wacc, 1.0)


def apply_dirichlet(A, b, fixed, values):
    """Symmetric elimination (fieldspice/operators.py:362-407):
    ``b -= A x0``; fixed rows/cols -> identity; ``b[fixed] = values``."""
    A = sp.csr_matrix(A, copy=True)
    b = np.array(b, dtype=float, copy=True)
    n = A.shape[0]
    fixed = np.asarray(fixed)
    idx = np.flatnonzero(fixed) if fixed.dtype == bool else fixed.astype(np.intp)
    vals = np.broadcast_to(np.asarray(values, dtype=float), idx.shape)
    x0 = np.zeros(n)
    x0[idx] = vals
    b -= A @ x0
    keep = np.ones(n)
    keep[idx] = 0.0
    Dm = sp.diags(keep)
    A = sp.csr_matrix(Dm @ A @ Dm + sp.diags(1.0 - keep))
    b[idx] = vals
    return A, b


def unique_last(idx, val):
    """(fieldspice/solvers/poisson.py:722-727)"""
    idx = np.asarray(idx)
    val = np.asarray(val, dtype=float)
    if idx.size == 0:
        return idx.astype(np.intp), val
    u, first_rev = np.unique(idx[::-1], return_index=True)
    return u.astype(np.intp), val[::-1][first_rev]


# --------------------------------------------------------------------------
# linear-algebra backend (fieldspice/solvers/poisson.py:159-589)
# --------------------------------------------------------------------------
class OracleConvergenceError(RuntimeError):
    def __init__(self, msg, history=()):
        super().__init__(msg)
        self.history = list(history)


def _row_abs_sum(A):
    n = A.shape[0]
    if A.nnz == 0:
        return np.zeros(n)
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    return np.bincount(rows, weights=np.abs(A.data), minlength=n)


class LinearSystem:
    """Restatement of ``LinearSystem`` (fieldspice/solvers/poisson.py:260-589)
    for symmetric matrices.

    ``method``: "direct" (SuperLU, MMD_AT_PLUS_A), "cg" (SciPy CG on the
    Jacobi-equilibrated free block with the reference's ILU preconditioner,
    falling back to Jacobi exactly as poisson.py:473-484 does) or "cg_jacobi"
    (the Jacobi branch forced: identity preconditioner on the equilibrated
    system -- the algorithm the GPU library runs).
    """

    def __init__(self, A, method="direct", tol=1e-10, maxiter=None, pin=True):
        t0 = time.perf_counter()
        A = sp.csr_matrix(A)
        n = A.shape[0]
        self.n = n
        self.tol = tol
        scale = float(np.abs(A.data).max()) if A.nnz else 0.0
        if scale == 0.0:
            raise ValueError("system matrix is identically zero")
        row_abs = _row_abs_sum(A)
        empty = np.flatnonzero(row_abs == 0.0)
        if empty.size:
            warnings.warn(f"{empty.size} unknown(s) are coupled to nothing (empty "
                          "rows); they are pinned to zero.", RuntimeWarning, stacklevel=2)
        rowsum = np.abs(A @ np.ones(n))
        keep = row_abs > 0.0
        self.singular = bool(keep.any() and np.all(rowsum[keep] <= 1e-9 * row_abs[keep]))
        pin_idx = list(empty)
        self.gauge = None
        if self.singular:
            if not pin:
                raise ValueError("system is singular")
            d = np.abs(A.diagonal())
            d[~keep] = -1.0
            self.gauge = int(np.argmax(d))
            pin_idx.append(self.gauge)
            warnings.warn(f"singular all-Neumann system: node {self.gauge} is "
                          "pinned to 0 to fix the gauge.", RuntimeWarning, stacklevel=2)
        self.pinned = np.array(sorted(set(pin_idx)), dtype=np.intp)
        if self.pinned.size:
            m = np.ones(n)
            m[self.pinned] = 0.0
            A = sp.csr_matrix(sp.diags(m) @ A @ sp.diags(m) + sp.diags(1.0 - m))
        dfull = A.diagonal()
        solo = (_row_abs_sum(A) <= np.abs(dfull)) & (dfull != 0)
        self.solo = np.flatnonzero(solo)
        self.free = np.flatnonzero(~solo)
        self.solo_diag = dfull[self.solo]
        self.split = bool(self.solo.size) and self.free.size > 0
        if self.split:
            A = sp.csr_matrix(A[self.free][:, self.free])
        d = A.diagonal()
        self.dhalf = np.sqrt(d)
        self.dinv = 1.0 / self.dhalf
        S = sp.diags(self.dinv)
        self.A = sp.csr_matrix(S @ A @ S)
        self.n_free = self.A.shape[0]
        self.method = method
        self.maxiter = maxiter
        self.prec = None
        if method == "direct":
            self.lu = spla.splu(self.A.tocsc(), permc_spec="MMD_AT_PLUS_A")
        elif method == "cg":
            try:
                ilu = spla.spilu(self.A.tocsc(), drop_tol=1e-4, fill_factor=10.0)
                self.M = spla.LinearOperator(self.A.shape, ilu.solve)
                self.prec = "ilu"
            except Exception:
                self.M = None
                self.prec = "jacobi"
        elif method == "cg_jacobi":
            self.M = None
            self.prec = "jacobi"
        else:
            raise ValueError(method)
        self.info = {"method": method, "n": n, "n_factorised": self.n_free,
                     "singular": self.singular, "pinned": self.gauge,
                     "setup_time": time.perf_counter() - t0}

    def solve(self, b, x0=None):
        t0 = time.perf_counter()
        b = np.array(b, dtype=float, copy=True)
        b[self.pinned] = 0.0
        nb = float(np.linalg.norm(b))
        bf = b[self.free] if self.split else b
        bs = bf * self.dinv
        iters = 0
        if self.method == "direct":
            # two steps of FP64 iterative refinement bring the backward error
            # of the LU solve to roundoff on ill-conditioned (kappa ~ 1e7)
            # systems, so the oracle is as exact as FP64 allows
            xs = self.lu.solve(bs)
            for _ in range(2):
                xs = xs + self.lu.solve(bs - self.A @ xs)
        else:
            maxiter = self.maxiter or max(500, int(10 * np.sqrt(self.n_free)))
            cnt = [0]

            def cb(_):
                cnt[0] += 1
            g0 = None
            if x0 is not None:
                g0 = np.asarray(x0, dtype=float)
                g0 = (g0[self.free] if self.split else g0) * self.dhalf
            xs, flag = spla.cg(self.A, bs, x0=g0, rtol=self.tol, atol=0.0,
                               maxiter=maxiter, M=self.M, callback=cb)
            iters = cnt[0]
        xf = xs * self.dinv
        if self.split:
            x = np.zeros(self.n)
            x[self.free] = xf
            x[self.solo] = b[self.solo] / self.solo_diag
        else:
            x = xf
        res_vec = (bs - self.A @ xs) * self.dhalf
        res = float(np.linalg.norm(res_vec) / nb) if nb > 0 else 0.0
        self.info.update(iterations=iters, residual=res,
                         residual_eq=float(np.linalg.norm(bs - self.A @ xs)
                                           / max(np.linalg.norm(bs), 1e-300)),
                         solve_time=time.perf_counter() - t0)
        if self.method != "direct":
            if res > max(1e-6, 100.0 * self.tol):
                raise OracleConvergenceError(
                    f"cg failed: relative residual {res:.3e} after {iters} iterations",
                    history=[res])
        return x


# --------------------------------------------------------------------------
# capacitance extraction (fieldspice/solvers/poisson.py:1185-1249)
# --------------------------------------------------------------------------
def capacitance_matrix(grid, eps_cell, terminal_nodes, wall_fixed=None,
                       method="direct", tol=1e-10, eps_mode="parallel",
                       return_phi=False):
    """Maxwell C by unit excitation; ``terminal_nodes`` is a list of node-index
    arrays; ``wall_fixed`` the Dirichlet-wall node indices (grounded shield).
    Returns (C symmetrised, report dict[, list of phi])."""
    L = nodal_laplacian(grid, cell_to_edge(grid, eps_cell, eps_mode))
    wall_idx = (np.empty(0, dtype=np.intp) if wall_fixed is None
                else np.asarray(wall_fixed, dtype=np.intp))
    term_nodes = np.concatenate(terminal_nodes)
    fixed, _ = unique_last(np.concatenate([wall_idx, term_nodes]),
                           np.zeros(wall_idx.size + term_nodes.size))
    shield = np.setdiff1d(wall_idx, term_nodes)
    A_bc, _ = apply_dirichlet(L, np.zeros(grid.n_nodes), fixed, np.zeros(fixed.size))
    ls = LinearSystem(A_bc, method=method, tol=tol)
    nt = len(terminal_nodes)
    C = np.zeros((nt, nt))
    qsh = np.zeros(nt)
    phis = []
    for j, tj in enumerate(terminal_nodes):
        x0 = np.zeros(grid.n_nodes)
        x0[tj] = 1.0
        b = -(L @ x0)
        b[fixed] = x0[fixed]
        phi = ls.solve(b)
        q = L @ phi
        for i, ti in enumerate(terminal_nodes):
            C[i, j] = float(q[ti].sum())
        if shield.size:
            qsh[j] = float(q[shield].sum())
        if return_phi:
            phis.append(phi)
    scale = float(np.abs(np.diag(C)).max()) or 1.0
    rep = {"asymmetry_abs": float(np.abs(C - C.T).max())}
    rep["asymmetry_rel"] = rep["asymmetry_abs"] / scale
    C = 0.5 * (C + C.T)
    rep["row_sum"] = C.sum(axis=1)
    rep["row_sum_rel"] = float(np.abs(rep["row_sum"]).max() / scale)
    off = C - np.diag(np.diag(C))
    rep["max_offdiag"] = float(off.max()) if nt > 1 else 0.0
    rep["shield_charge"] = qsh
    rep["linear"] = dict(ls.info)
    if return_phi:
        return C, rep, phis
    return C, rep


# --------------------------------------------------------------------------
# EQS transient (fieldspice/solvers/eqs.py:176-195, 229-375)
# --------------------------------------------------------------------------
def eqs_transient(grid, eps_cell, sigma_cell, fixed_nodes_fn, t_end, dt,
                  theta=1.0, t_start=0.0, method="direct", tol=1e-10,
                  terminal_nodes=(), phi0=None, keep_phis=True):
    """Backward-Euler / theta EQS.  ``fixed_nodes_fn(t) -> (idx, vals)`` gives
    the Dirichlet set at time t (voltage-driven terminals and walls, already
    merged with the later-wins rule).  Returns dict with times, phis (list,
    or only the last one), terminal currents (per terminal_nodes entry)."""
    Le = nodal_laplacian(grid, cell_to_edge(grid, eps_cell))
    Ls = nodal_laplacian(grid, cell_to_edge(grid, sigma_cell))
    n = grid.n_nodes
    if phi0 is None:
        idx, vals = fixed_nodes_fn(t_start)
        if idx.size == 0:
            phi = np.zeros(n)
        else:
            A, b = apply_dirichlet(Le, np.zeros(n), idx, vals)
            phi = LinearSystem(A, method=method, tol=tol).solve(b)
    else:
        phi = np.asarray(phi0, dtype=float).copy()
    nsteps = int(np.ceil((t_end - t_start) / dt))
    times = t_start + dt * np.arange(nsteps + 1)
    K = sp.csr_matrix(Le / dt + theta * Ls)
    Kexp = sp.csr_matrix(Le / dt - (1.0 - theta) * Ls)
    prepared = {}
    phis = [phi.copy()]
    currents = [np.zeros(nsteps + 1) for _ in terminal_nodes]
    prev = phi
    for k, t in enumerate(times[1:], start=1):
        idx, vals = fixed_nodes_fn(t)
        rhs = Kexp @ prev
        A, b = apply_dirichlet(K, rhs, idx, vals)
        key = np.asarray(idx).tobytes()
        if key not in prepared:
            prepared[key] = LinearSystem(A, method=method, tol=tol)
        phi = prepared[key].solve(b)
        # current into the mesh, evaluated with theta = 1 as the reference's
        # _record_terminals does (eqs.py:328-340,360-375)
        resid = Le @ (phi - prev) / dt + Ls @ phi
        for m, tn in enumerate(terminal_nodes):
            currents[m][k] = float(resid[tn].sum())
        if keep_phis:
            phis.append(phi.copy())
        prev = phi
    for c in currents:
        if c.size > 1:
            c[0] = c[1]
    return {"t": times, "phis": phis if keep_phis else [prev],
            "currents": currents, "n_factorisations": len(prepared)}


# --------------------------------------------------------------------------
# thermal (fieldspice/solvers/thermal.py:140-293)
# --------------------------------------------------------------------------
def thermal_matrices(grid, kappa_cell, rho_cp_cell=None, conv_diag=None):
    K = nodal_laplacian(grid, cell_to_edge(grid, kappa_cell))
    if conv_diag is not None:
        K = sp.csr_matrix(K + sp.diags(conv_diag))
    Cth = None
    if rho_cp_cell is not None:
        vol = grid.node_volumes().ravel()
        Cth = cell_to_node(grid, rho_cp_cell).ravel() * vol
    return K, Cth


def thermal_steady(grid, kappa_cell, q_node, fixed=None, vals=None,
                   conv_diag=None, conv_rhs=None, method="direct", tol=1e-10):
    K, _ = thermal_matrices(grid, kappa_cell, None, conv_diag)
    b = np.asarray(q_node, dtype=float).copy()
    if conv_rhs is not None:
        b = b + conv_rhs
    if fixed is not None and np.size(fixed):
        K, b = apply_dirichlet(K, b, fixed, vals)
    return LinearSystem(K, method=method, tol=tol).solve(b)


def thermal_transient(grid, kappa_cell, rho_cp_cell, q_fn, t_end, dt, T0,
                      fixed_fn=None, conv_diag=None, conv_rhs=None,
                      theta=1.0, method="direct", tol=1e-10):
    K, Cth = thermal_matrices(grid, kappa_cell, rho_cp_cell, conv_diag)
    n = grid.n_nodes
    T = np.full(n, float(T0)) if np.isscalar(T0) else np.asarray(T0, float).copy()
    C = sp.diags(Cth / dt)
    A_imp = sp.csr_matrix(C + theta * K)
    A_exp = sp.csr_matrix(C - (1.0 - theta) * K)
    nsteps = int(np.ceil(t_end / dt))
    times = dt * np.arange(nsteps + 1)
    wall = np.zeros(n) if conv_rhs is None else conv_rhs
    Ts = [T.copy()]
    prepared = {}
    for t in times[1:]:
        q = np.asarray(q_fn(t), dtype=float).ravel()
        if q.size == 1:
            q = np.full(n, float(q))
        b = A_exp @ T + q + wall
        idx, vals = fixed_fn(t) if fixed_fn is not None else (np.empty(0, np.intp), np.empty(0))
        A, bb = apply_dirichlet(A_imp, b, idx, vals) if np.size(idx) else (A_imp, b)
        key = np.asarray(idx).tobytes()
        if key not in prepared:
            prepared[key] = LinearSystem(A, method=method, tol=tol)
        T = prepared[key].solve(bb)
        Ts.append(T.copy())
    return times, Ts


# --------------------------------------------------------------------------
# island coarse space of high-contrast systems (no reference counterpart: the
# library's own definition, restated here as the checker; pinned by a
# constructed case in tests/test_oracle.py)
# --------------------------------------------------------------------------
def floating_islands(grid, w_edge, anchored, contrast=1e3):
    """Components of the edges with ``w > sqrt(w_min*w_max)`` (positive
    weights) holding >= 2 nodes and no anchored (Dirichlet or pinned) node,
    when ``w_max >= contrast*w_min``.  Returns (labels, count): labels per node
    (C order), -1 outside an island, islands numbered by their smallest node
    in x-fastest order (i + nx*(j + ny*k))."""
    from scipy.sparse.csgraph import connected_components
    n = grid.n_nodes
    w = np.asarray(w_edge, dtype=float)
    pos = w[w > 0]
    labels = np.full(n, -1, np.int64)
    if pos.size == 0 or pos.max() < contrast * pos.min():
        return labels, 0
    thr = np.sqrt(pos.min() * pos.max())
    G = grad_node_edge(grid).tocoo()
    tail = np.empty(G.shape[0], np.int64)
    head = np.empty(G.shape[0], np.int64)
    tail[G.row[G.data < 0]] = G.col[G.data < 0]
    head[G.row[G.data > 0]] = G.col[G.data > 0]
    s = w > thr
    S = sp.coo_matrix((np.ones(int(s.sum())), (tail[s], head[s])), shape=(n, n))
    _, comp = connected_components(S, directed=False)
    size = np.bincount(comp)
    anch = np.zeros(size.size, bool)
    anch[comp[np.asarray(anchored, bool)]] = True
    keep = (size >= 2) & ~anch
    nx, ny, nz = grid.shape_nodes
    i, j, k = np.unravel_index(np.arange(n), (nx, ny, nz))
    xfast = i + nx * (j + ny * k)
    first = np.full(size.size, np.iinfo(np.int64).max)
    np.minimum.at(first, comp, xfast)
    ids = np.nonzero(keep)[0]
    ids = ids[np.argsort(first[ids])]
    remap = np.full(size.size, -1, np.int64)
    remap[ids] = np.arange(ids.size)
    return remap[comp], int(ids.size)


def island_matrix(L_free, labels, count):
    """E = 1_I^T L_ff 1_J for the masked (Dirichlet rows and columns removed)
    operator ``L_free``."""
    n = L_free.shape[0]
    rows = np.nonzero(labels >= 0)[0]
    Z = sp.csr_matrix((np.ones(rows.size), (rows, labels[rows])), shape=(n, count))
    return (Z.T @ L_free @ Z).toarray()


# --------------------------------------------------------------------------
# nonlinear Poisson (fieldspice/solvers/poisson.py:1380-1524, 1550-1617)
# --------------------------------------------------------------------------
Q_ELEM = 1.602_176_634e-19
KB = 1.380_649e-23


def boltzmann_charge(doping_node, ni, T=300.0):
    """rho = q (p - n + N) with n, p = ni exp(+-phi/Vt), exponent clipped at
    400; drho = -q (n + p)/Vt; guess Vt asinh(N/2ni)
    (fieldspice/solvers/poisson.py:1550-1617).  Returns (rho, drho, guess, Vt)."""
    dop = np.asarray(doping_node, dtype=float).ravel()
    Vt = KB * T / Q_ELEM

    def carriers(phi):
        z = np.clip(np.asarray(phi, dtype=float) / Vt, -400.0, 400.0)
        return ni * np.exp(z), ni * np.exp(-z)

    def rho(phi):
        n, p = carriers(phi)
        return Q_ELEM * (p - n + dop)

    def drho(phi):
        n, p = carriers(phi)
        return -Q_ELEM * (n + p) / Vt

    return rho, drho, Vt * np.arcsinh(dop / (2.0 * ni)), Vt


def nonlinear_poisson(grid, eps_cell, rho, drho, fixed=None, vals=None, rho_node=None,
                      phi0=None, v_scale=None, step_limit=5.0, armijo=1e-4, newton_tol=1e-8,
                      max_newton=60, damping=1.0, floating=(), charges=(), merge=()):
    """Damped Newton for L phi - rho(phi) V_node = rho_node V_node with
    Dirichlet rows ``fixed`` at ``vals`` (fieldspice/solvers/poisson.py:1429-1511):
    the Jacobian L - diag(drho V_node) is eliminated and solved by SuperLU each
    step, the first trial step is clamped to step_limit*v_scale, then Armijo
    backtracking.  Node sets in ``floating`` are merged into one unknown each,
    carrying ``charges`` (the reference's reduction, :1432-1451), and so are
    the sets in ``merge`` (periodic wall pairs).  Returns (phi, history of
    |F|, damping factors)."""
    L = nodal_laplacian(grid, cell_to_edge(grid, eps_cell))
    vnode = np.asarray(grid.node_volumes(), dtype=float).ravel()
    n = grid.n_nodes
    fixed = np.empty(0, dtype=np.intp) if fixed is None else np.asarray(fixed, dtype=np.intp)
    vals = np.empty(0) if vals is None else np.asarray(vals, dtype=float)
    b_ext = np.zeros(n) if rho_node is None else np.asarray(rho_node, float).ravel() * vnode
    b_ext = b_ext.copy()
    for g, q in zip(floating, charges):
        b_ext[np.asarray(g, dtype=np.intp)[0]] += float(q)
    group, P = merge_map(n, list(floating) + list(merge))
    fixed_r = group[fixed]
    v_scale = KB * 300.0 / Q_ELEM if v_scale is None else v_scale
    phi = np.zeros(n) if phi0 is None else np.asarray(phi0, dtype=float).ravel().copy()
    phi[fixed] = vals

    def residual(p):
        f = P.T @ (L @ p - rho(p) * vnode - b_ext)
        f[fixed_r] = 0.0
        return f

    F = residual(phi)
    hist = [float(np.linalg.norm(F))]
    lams = []
    for it in range(max_newton):
        if hist[-1] <= 1e-14 * max(hist[0], 1e-300):
            return phi, hist, lams
        J = sp.csr_matrix(P.T @ (L - sp.diags(drho(phi) * vnode)) @ P)
        A, b = apply_dirichlet(J, -F, fixed_r, np.zeros(fixed_r.size))
        dphi = P @ spla.splu(A.tocsc()).solve(b)
        big = float(np.abs(dphi).max())
        lam = float(damping)
        if big > 0.0:
            lam = min(lam, step_limit * v_scale / big)
        f0 = hist[-1]
        while True:
            Fn = residual(phi + lam * dphi)
            fn = float(np.linalg.norm(Fn))
            if np.isfinite(fn) and fn < (1.0 - armijo * lam) * f0:
                break
            lam *= 0.5
            if lam < 1e-12:
                raise OracleConvergenceError("line search collapsed")
        phi = phi + lam * dphi
        F = Fn
        hist.append(fn)
        lams.append(lam)
        if lam * big <= newton_tol * v_scale:
            return phi, hist, lams
    raise OracleConvergenceError("Newton did not converge")


# --------------------------------------------------------------------------
# field region AC admittance (fieldspice/circuit/coupling.py:216-234)
# --------------------------------------------------------------------------
def field_region_ac(grid, eps_cell, sigma_cell, terminal_nodes, omega):
    """``Y = K_tt - K_ti K_ii^-1 K_it`` of ``K = L_sigma + j omega L_eps``
    with ``K_tt``/``K_ti`` the electrode sums of ``K`` and ``K_ii`` the
    interior block, factorised by complex ``splu`` as the reference does."""
    Ls = nodal_laplacian(grid, cell_to_edge(grid, sigma_cell, "parallel"))
    Le = nodal_laplacian(grid, cell_to_edge(grid, eps_cell, "parallel"))
    K = (Ls.astype(complex) + 1j * float(omega) * Le).tocsr()
    n = grid.n_nodes
    nt = len(terminal_nodes)
    rows, cols = [], []
    for k, t in enumerate(terminal_nodes):
        rows.append(np.full(len(t), k))
        cols.append(np.asarray(t, dtype=np.intp))
    P = sp.csr_matrix((np.ones(sum(len(t) for t in terminal_nodes)),
                       (np.concatenate(rows), np.concatenate(cols))), shape=(nt, n))
    interior = np.setdiff1d(np.arange(n), np.concatenate(cols))
    S = sp.csr_matrix((np.ones(interior.size), (np.arange(interior.size), interior)),
                      shape=(interior.size, n))
    K_tt = (P @ K @ P.T).toarray()
    if interior.size == 0:
        return K_tt
    K_ti = (P @ K @ S.T).tocsr()
    lu = spla.splu((S @ K @ S.T).tocsc())
    W = lu.solve(np.asarray((K_ti.T).todense()))
    return K_tt - K_ti @ W


# --------------------------------------------------------------------------
# Poisson with floating / charge-driven electrodes
# (fieldspice/solvers/poisson.py:1039-1063, node merge :663-720)
# --------------------------------------------------------------------------
def poisson_floating(grid, eps_cell, fixed, vals, floating, charges, b_free=None,
                     merge=(), return_info=False):
    """The reference's node merge (fieldspice/solvers/poisson.py:665-719,
    1030-1063): each node set in ``floating`` and in ``merge`` (periodic wall
    pairs) becomes one unknown (``phi = P x``, ``P^T L P x = P^T b``), reduced
    indices in the order of each set's smallest node (as scipy's
    connected_components numbers them); each floating set's prescribed charge
    is added to its merged row; the Dirichlet nodes ``fixed`` are eliminated
    on the reduced indices (a merged unknown with several Dirichlet nodes
    takes the last one's value); solved by ``LinearSystem`` ("direct"), which
    pins the gauge of a singular system.  Returns the node potential (and the
    solver's info)."""
    L = nodal_laplacian(grid, cell_to_edge(grid, eps_cell, "parallel"))
    n = grid.n_nodes
    group, P = merge_map(n, list(floating) + list(merge))
    b = np.zeros(n) if b_free is None else np.asarray(b_free, dtype=float).copy()
    for g, q in zip(floating, charges):
        b[np.asarray(g, dtype=np.intp)[0]] += float(q)
    A_r = sp.csr_matrix(P.T @ L @ P)
    b_r = P.T @ b
    A_bc, b_bc = apply_dirichlet(A_r, b_r, group[np.asarray(fixed, dtype=np.intp)],
                                 np.asarray(vals, dtype=float))
    ls = LinearSystem(A_bc, method="direct")
    phi = P @ ls.solve(b_bc)
    if return_info:
        info = dict(ls.info)
        info["group"] = group
        return phi, info
    return phi


def merge_map(n, sets):
    """Node merge of ``fieldspice/solvers/poisson.py:683-711``: the union of
    the node sets as connected components, numbered in the order of each
    component's smallest node; returns (reduced index per node, P)."""
    label = np.arange(n)
    sets = [np.asarray(g, dtype=np.intp) for g in sets]
    changed = True
    while changed:                       # union of overlapping sets
        changed = False
        for g in sets:
            lo = label[g].min()
            if np.any(label[g] != lo):
                label[np.isin(label, label[g])] = lo
                changed = True
    _, group = np.unique(label, return_inverse=True)
    nred = int(group.max()) + 1
    return group, sp.csr_matrix((np.ones(n), (np.arange(n), group)), shape=(n, nred))


def periodic_pairs(grid, axis):
    """(lo, hi) node indices of the two walls normal to ``axis``, paired
    position by position (``fieldspice/boundaries.py:226-268``)."""
    nid = np.arange(grid.n_nodes).reshape(grid.shape_nodes)
    return (np.take(nid, 0, axis=axis).ravel(), np.take(nid, -1, axis=axis).ravel())


def periodic_sets(grid, axes):
    """Merge sets identifying the wall pairs of every axis in ``axes``."""
    out = []
    for a in axes:
        lo, hi = periodic_pairs(grid, a)
        out.extend(np.array([x, y]) for x, y in zip(lo, hi))
    return out


def capacitance_matrix_merged(grid, eps_cell, terminal_nodes, merge, wall_fixed=None):
    """Maxwell C with node sets merged (periodic wall pairs), as
    ``fieldspice/solvers/poisson.py:1185-1249`` runs it: ``P^T L P`` with
    every unknown holding a terminal or wall node fixed, excitation j setting
    the unknowns of terminal j's nodes to 1, charges ``L phi`` summed over each
    terminal's own nodes on the full grid.  Solved directly."""
    L = nodal_laplacian(grid, cell_to_edge(grid, eps_cell, "parallel"))
    n = grid.n_nodes
    group, P = merge_map(n, merge)
    A_r = sp.csr_matrix(P.T @ L @ P)
    wall_idx = (np.empty(0, dtype=np.intp) if wall_fixed is None
                else np.asarray(wall_fixed, dtype=np.intp))
    term_nodes = np.concatenate(terminal_nodes)
    fixed, _ = unique_last(np.concatenate([wall_idx, term_nodes]),
                           np.zeros(wall_idx.size + term_nodes.size))
    fixed_r = group[fixed]
    shield = np.setdiff1d(wall_idx, term_nodes)
    A_bc, _ = apply_dirichlet(A_r, np.zeros(A_r.shape[0]), fixed_r, np.zeros(fixed_r.size))
    ls = LinearSystem(A_bc, method="direct")
    nt = len(terminal_nodes)
    C = np.zeros((nt, nt))
    qsh = np.zeros(nt)
    for j, tj in enumerate(terminal_nodes):
        x0 = np.zeros(A_r.shape[0])
        x0[group[tj]] = 1.0
        b = -(A_r @ x0)
        b[fixed_r] = x0[fixed_r]
        q = L @ (P @ ls.solve(b))
        for i, ti in enumerate(terminal_nodes):
            C[i, j] = float(q[ti].sum())
        if shield.size:
            qsh[j] = float(q[shield].sum())
    return 0.5 * (C + C.T), qsh
