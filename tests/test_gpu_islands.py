"""Island coarse space of high-contrast systems (copper islands floating in
SiO2 under an EQS step) on the GPU: island detection and the coarse matrix
against the oracle, deflated FP64 and mixed-precision solves against the
direct solve, and the EQS transient against the oracle's direct time loop."""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import fs_oracle as O
from fieldspice_amd import _lib
from fieldspice_amd.solvers.base import SolverConfig
from fieldspice_amd.solvers.eqs import EQSSolver
from fieldspice_amd.workloads import eqs_two_phase

pytestmark = pytest.mark.gpu

DT = 1e-12


def _case(n=24, seed=2):
    g, eps, sig, terms = eqs_two_phase(n, seed=seed)
    w = (O.edge_weights(g, O.cell_to_edge(g, eps)) / DT
         + O.edge_weights(g, O.cell_to_edge(g, sig)))
    fixed = np.zeros(g.n_nodes, bool)
    for t in terms:
        fixed[t.nodes] = True
    op = _lib.DeviceOperator.from_edges(g, w)
    sysd = _lib.DeviceSystem(op, fixed.astype(np.uint8))
    L = O.nodal_laplacian(g, O.cell_to_edge(g, eps)) / DT + O.nodal_laplacian(
        g, O.cell_to_edge(g, sig))
    return g, w, fixed, op, sysd, L.tocsr(), terms


def _masked(L, fixed):
    """The eliminated operator: L_ff on free rows, identity on fixed rows."""
    free = ~fixed
    return (sp.diags(free * 1.0) @ L @ sp.diags(free * 1.0) + sp.diags(fixed * 1.0)).tocsc()


def _rhs(L, fixed, terms):
    xD = np.zeros(L.shape[0])
    xD[terms[0].nodes] = 1.0
    b = -(L @ xD)
    b[fixed] = xD[fixed]
    return b


def test_islands_and_coarse_matrix_match_oracle():
    print("[gpu case] island detection and E vs oracle", flush=True)
    g, w, fixed, op, sysd, L, terms = _case()
    ids, n = sysd.find_islands(1e3)
    ref, nref = O.floating_islands(g, w, fixed)
    assert n == nref == 5
    np.testing.assert_array_equal(ids, ref)    # same numbering (smallest x-fast node)
    sysd.set_deflation(ids, n)
    m, einv = sysd.deflation()
    free = ~fixed
    Lf = (sp.diags(free * 1.0) @ L @ sp.diags(free * 1.0)).tocsr()
    E = O.island_matrix(Lf, ref, nref)
    # E^-1 from the library's Cholesky factorisation, E from the oracle's CSR
    # product: equal up to the conditioning of E (islands one gap apart couple)
    np.testing.assert_allclose(einv @ E, np.eye(n), atol=1e-10)
    np.testing.assert_allclose(einv, np.linalg.inv(E), rtol=1e-8, atol=0)
    sysd.set_deflation()                       # removal
    assert sysd.deflation()[0] == 0
    # no islands below the contrast threshold
    _, n_hi = sysd.find_islands(1e12)
    assert n_hi == 0


@pytest.mark.parametrize("precision", [_lib.FS_PREC_FP64, _lib.FS_PREC_MIXED])
def test_deflated_solve_matches_direct(precision):
    print(f"[gpu case] deflated solve precision={precision}", flush=True)
    g, w, fixed, op, sysd, L, terms = _case()
    b = _rhs(L, fixed, terms)
    xref = spla.spsolve(_masked(L, fixed), b)
    # the reference's method (Jacobi-PCG) at tol 1e-10: its distance to the
    # exact solution is set by the conditioning of this system
    x0, st0, rc0 = sysd.solve(b, _lib.DeviceSystem.params(1e-10, maxiter=3000,
                                                          precision=_lib.FS_PREC_FP64))
    err0 = np.abs(x0 - xref).max() / np.abs(xref).max()
    ids, n = sysd.find_islands()
    sysd.set_deflation(ids, n)
    x, st, rc = sysd.solve(b, _lib.DeviceSystem.params(1e-10, precision=precision))
    err = np.abs(x - xref).max() / np.abs(xref).max()
    print(f"plain FP64: {st0['iterations']} its (rc {rc0}), err {err0:.1e}; with islands: "
          f"{st['iterations']} its, {st['refreshes']} refreshes, residual_eq "
          f"{st['residual_eq']:.2e}, err {err:.1e}", flush=True)
    assert rc == 0 and st["converged"] == 1
    assert st["residual_eq"] < 1e-10 * 1.01
    # FP32 cycles restart at each FP64 refresh: fewer savings, same answer
    assert st["iterations"] < st0["iterations"] / (2 if precision == _lib.FS_PREC_FP64 else 1)
    assert err <= 3 * err0 + 1e-12               # the same accuracy class as the reference
    # tighter tolerance: the answer is the direct solve's
    x, st, rc = sysd.solve(b, _lib.DeviceSystem.params(1e-13, precision=precision))
    assert rc == 0
    assert np.abs(x - xref).max() <= 1e-9 * np.abs(xref).max()


def test_deflated_multi_rhs_and_argument_checks():
    print("[gpu case] deflated multi-RHS, argument checks", flush=True)
    g, w, fixed, op, sysd, L, terms = _case(n=20, seed=0)
    ids, n = sysd.find_islands()
    assert n == 4
    rng = np.random.default_rng(3)
    B = rng.normal(size=(3, L.shape[0])) * (~fixed)
    sysd.set_deflation(ids, n)
    Lf = _masked(L, fixed)
    dh = np.where(fixed, 0.0, 1.0 / np.sqrt(L.diagonal()))
    for prec in (_lib.FS_PREC_FP64, _lib.FS_PREC_MIXED):
        X, st, rc = sysd.solve(B, _lib.DeviceSystem.params(1e-11, precision=prec))
        assert rc == 0
        for k in range(3):
            # the stop rule recomputed from the oracle's CSR: equilibrated residual
            res = np.linalg.norm(dh * (B[k] - Lf @ X[k])) / np.linalg.norm(dh * B[k])
            assert res < 1.05e-11, (prec, k, res)
            ref = spla.spsolve(Lf, B[k])
            assert np.abs(X[k] - ref).max() <= 1e-6 * np.abs(ref).max(), (prec, k)
    bad = ids.copy()
    bad[np.nonzero(fixed)[0][0]] = 0              # a Dirichlet node inside an island
    with pytest.raises(_lib.LibraryError):
        sysd.set_deflation(bad, n)
    with pytest.raises(_lib.LibraryError):
        sysd.set_deflation(ids, 1025)                # above the 1024-island limit
    with pytest.raises(_lib.LibraryError):      # no FCG with an island space
        sysd.solve(B[0], _lib.DeviceSystem.params(1e-10, precision=_lib.FS_PREC_FCG))


def test_eqs_transient_with_islands_matches_oracle():
    print("[gpu case] EQS transient, islands, FP64 and mixed vs oracle", flush=True)
    g, eps, sig, terms = eqs_two_phase(20, seed=0)
    nsteps = 6

    def fixed_fn(t):
        idx = np.concatenate([terms[0].nodes, terms[1].nodes])
        vals = np.concatenate([np.full(terms[0].nodes.size, terms[0].value_at(t)),
                               np.zeros(terms[1].nodes.size)])
        return idx, vals
    ref = O.eqs_transient(g, eps, sig, fixed_fn, nsteps * DT, DT)
    phis = {}
    for name in ("hip_pcg", "hip_mixed"):
        # tol 1e-13: the steps are ill-conditioned (copper vs SiO2 ~ 1e6), so
        # the distance to the direct solve is ~ cond * tol
        cfg = SolverConfig(linear_solver=name, deflate_islands=True, tol=1e-13)
        res = EQSSolver(g, eps, sig, cfg).solve(terms, nsteps * DT, DT, store=True)
        assert res.meta["linear"]["islands"] == 4
        phis[name] = res.fields["phi"]
    for k in range(1, nsteps + 1):
        r = ref["phis"][k]
        scale = np.abs(r).max()
        for name, ph in phis.items():
            err = np.abs(ph[k] - r).max() / scale
            assert err < 1e-9, (name, k, err)


def test_thermal_transient_with_islands_matches_oracle():
    # copper (400 W/mK) boxes in SiO2 (1.4 W/mK), heat into the driven box, the
    # x = 0 wall held at 300 K: islands carry the C/dt node diagonal in E and in
    # the FP32 difference-form operator
    print("[gpu case] thermal transient, islands, FP64 and mixed vs oracle", flush=True)
    from fieldspice_amd.boundaries import BoundarySpec, Dirichlet
    from fieldspice_amd.grid import RectilinearGrid
    from fieldspice_amd.solvers.thermal import ThermalSolver
    g0, _, sig, terms = eqs_two_phase(20, seed=0)
    g = RectilinearGrid(g0.ax.nodes * 100.0, g0.ay.nodes * 100.0, g0.az.nodes * 100.0)
    kappa = np.where(sig > 0, 400.0, 1.4)
    rcp = np.full(g.shape_cells, 2.0e6)
    dt, nsteps, T0 = 1e-5, 5, 300.0
    q0 = np.zeros(g.n_nodes)
    q0[terms[0].nodes] = 1e-3 / terms[0].nodes.size
    qfn = lambda t: q0          # noqa: E731
    wall = g.node_ids()[0].ravel()
    _, Ts = O.thermal_transient(g, kappa, rcp, qfn, nsteps * dt, dt, T0,
                                fixed_fn=lambda t: (wall, np.full(wall.size, T0)))
    Tref = np.array(Ts)
    for name in ("hip_pcg", "hip_mixed"):
        cfg = SolverConfig(linear_solver=name, deflate_islands=True, island_contrast=100.0,
                           tol=1e-13)
        r = ThermalSolver(g, kappa, rcp, cfg).transient(qfn, nsteps * dt, dt, T0=T0,
                                                        bc=BoundarySpec(xlo=Dirichlet(T0)),
                                                        store=True)
        assert r.meta["linear"]["islands"] > 0
        err = np.abs(r.fields["T"] - Tref).max() / np.abs(Tref - T0).max()
        print(f"{name}: {r.meta['linear']['islands']} islands, max rel err {err:.2e}", flush=True)
        assert err < 1e-9, (name, err)
