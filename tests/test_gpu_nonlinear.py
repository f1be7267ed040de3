"""GPU nonlinear Poisson (Newton, Jacobian L + positive node diagonal) against
the pn-junction built-in potential and the oracle's Newton."""

import numpy as np
import pytest

import fs_oracle as O
from fieldspice_amd.grid import RectilinearGrid
from fieldspice_amd.solvers.base import ConvergenceError, SolverConfig, Terminal
from fieldspice_amd.solvers.poisson import NonlinearPoissonSolver, boltzmann_equilibrium_charge
from fieldspice_amd.units import eps0

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[gpu case] {msg}", flush=True)


def test_pn_junction_built_in_potential():
    _case("nonlinear Poisson: 1D Si pn junction, Vbi")
    L, N, ni = 1e-6, 1e23, 1e16
    g = RectilinearGrid(np.linspace(0.0, L, 1001), None, None)
    x = np.broadcast_to(g.xn[:, None, None], g.shape_nodes).ravel()
    dop = np.where(x < L / 2, -N, N)
    ec = boltzmann_equilibrium_charge(dop, ni)
    nid = g.node_ids()
    lo, hi = nid[0].ravel(), nid[-1].ravel()
    s = NonlinearPoissonSolver(g, np.full(g.shape_cells, 11.7 * eps0), ec.rho, ec.drho_dphi)
    r = s.solve([Terminal("p", lo, voltage=float(ec.phi_guess[lo][0]))], phi0=ec.phi_guess)
    phi = r.fields["phi"][0].ravel()
    vbi = float(phi[hi].mean() - phi[lo].mean())
    # docs/CONTRACTS.md:455,480-481 of the reference: Vt ln(Na Nd/ni^2) to 1e-9
    assert vbi == pytest.approx(ec.Vt * np.log(N * N / ni ** 2), rel=1e-9)
    assert vbi == pytest.approx(0.833370, abs=5e-7)
    nt = r.meta["newton"]
    assert 3 <= nt["iterations"] <= 40
    assert r.scalars["residual"][-1] < 1e-8 * r.scalars["residual"][0]
    # against the oracle's Newton with direct solves
    rho, drho, guess, _ = O.boltzmann_charge(dop, ni)
    po, _, _ = O.nonlinear_poisson(g, np.full(g.shape_cells, 11.7 * eps0), rho, drho,
                                   fixed=lo, vals=guess[lo], phi0=guess)
    assert np.abs(phi - po).max() <= 1e-9 * np.abs(po).max()
    assert r.fields["e"].shape == (1, sum(a.size for a in g.edge_lengths()))


def test_3d_well_matches_oracle_newton():
    _case("nonlinear Poisson: 3D n-well in p-substrate, two contacts, vs oracle")
    g = RectilinearGrid(np.linspace(0, 0.6e-6, 31), np.linspace(0, 0.4e-6, 21),
                        np.linspace(0, 0.3e-6, 16))
    xc = np.broadcast_to(g.xn[:, None, None], g.shape_nodes)
    yc = np.broadcast_to(g.yn[None, :, None], g.shape_nodes)
    zc = np.broadcast_to(g.zn[None, None, :], g.shape_nodes)
    well = (xc > 0.2e-6) & (xc < 0.45e-6) & (yc > 0.1e-6) & (yc < 0.3e-6) & (zc > 0.15e-6)
    dop = np.where(well, 5e22, -2e22).ravel()
    ni = 1e16
    ec = boltzmann_equilibrium_charge(dop, ni)
    nid = g.node_ids()
    sub = nid[:, :, 0].ravel()                                   # substrate contact
    wc = nid[15:19, 8:13, -1].ravel()                            # well contact, 0.3 V bias
    fixed = np.concatenate([sub, wc])
    vals = np.concatenate([ec.phi_guess[sub], ec.phi_guess[wc] + 0.3])
    eps = np.full(g.shape_cells, 11.7 * eps0)
    s = NonlinearPoissonSolver(g, eps, ec.rho, ec.drho_dphi, SolverConfig(tol=1e-12))
    r = s.solve([Terminal("sub", sub, voltage=float(vals[0])),
                 Terminal("well", wc, voltage=float(vals[-1]))], phi0=ec.phi_guess)
    phi = r.fields["phi"][0].ravel()
    rho, drho, guess, _ = O.boltzmann_charge(dop, ni)
    po, hist, _ = O.nonlinear_poisson(g, eps, rho, drho, fixed=fixed, vals=vals, phi0=guess)
    assert np.abs(phi - po).max() <= 1e-8 * np.abs(po).max()
    assert np.all(phi[wc] == pytest.approx(vals[-1]))
    # the well contact's charge equals the oracle's L phi - rho V on its nodes
    Lo = O.nodal_laplacian(g, O.cell_to_edge(g, eps))
    qo = Lo @ po - rho(po) * g.node_volumes().ravel()
    assert r.terminals["well"]["q"][0] == pytest.approx(qo[wc].sum(), rel=1e-6,
                                                         abs=1e-6 * np.abs(qo[wc]).sum())


def test_wrong_derivative_raises_convergence_error():
    _case("nonlinear Poisson: non-descent direction -> ConvergenceError")
    g = RectilinearGrid(np.linspace(0.0, 1e-6, 201), None, None)
    x = np.broadcast_to(g.xn[:, None, None], g.shape_nodes).ravel()
    dop = np.where(x < 0.5e-6, -1e23, 1e23)
    ec = boltzmann_equilibrium_charge(dop, 1e16)
    lo = g.node_ids()[0].ravel()
    s = NonlinearPoissonSolver(g, np.full(g.shape_cells, 11.7 * eps0), ec.rho,
                               lambda p: -ec.drho_dphi(p), SolverConfig(max_newton=30))
    with pytest.raises(ConvergenceError):
        s.solve([Terminal("p", lo, voltage=float(ec.phi_guess[lo][0]))], phi0=ec.phi_guess)


@pytest.mark.parametrize("Qw", [0.0, 2e-18])
def test_3d_well_with_floating_contact_matches_oracle_merge(Qw):
    _case(f"nonlinear Poisson: 3D n-well, floating well contact Q={Qw}, vs oracle merge")
    g = RectilinearGrid(np.linspace(0, 0.6e-6, 31), np.linspace(0, 0.4e-6, 21),
                        np.linspace(0, 0.3e-6, 16))
    xc = np.broadcast_to(g.xn[:, None, None], g.shape_nodes)
    yc = np.broadcast_to(g.yn[None, :, None], g.shape_nodes)
    zc = np.broadcast_to(g.zn[None, None, :], g.shape_nodes)
    well = (xc > 0.2e-6) & (xc < 0.45e-6) & (yc > 0.1e-6) & (yc < 0.3e-6) & (zc > 0.15e-6)
    dop = np.where(well, 5e22, -2e22).ravel()
    ec = boltzmann_equilibrium_charge(dop, 1e16)
    nid = g.node_ids()
    sub = nid[:, :, 0].ravel()
    wc = nid[15:19, 8:13, -1].ravel()
    guess = ec.phi_guess.copy()
    guess[wc] = guess[wc].mean()                 # the contact starts as one equipotential
    eps = np.full(g.shape_cells, 11.7 * eps0)
    s = NonlinearPoissonSolver(g, eps, ec.rho, ec.drho_dphi, SolverConfig(tol=1e-12))
    r = s.solve([Terminal("sub", sub, voltage=float(guess[sub][0])),
                 Terminal("well", wc, current=(Qw or None))], phi0=guess)
    phi = r.fields["phi"][0].ravel()
    rho, drho, _, _ = O.boltzmann_charge(dop, 1e16)
    po, _, _ = O.nonlinear_poisson(g, eps, rho, drho, fixed=sub, vals=guess[sub], phi0=guess,
                                   floating=[wc], charges=[Qw])
    assert np.abs(phi - po).max() <= 1e-8 * np.abs(po).max()
    assert np.ptp(phi[wc]) < 1e-12
    Lo = O.nodal_laplacian(g, O.cell_to_edge(g, eps))
    qo = Lo @ po - rho(po) * g.node_volumes().ravel()
    scale = np.abs(qo[wc]).sum()
    assert r.terminals["well"]["q"][0] == pytest.approx(Qw, abs=1e-6 * scale)
    assert r.meta["n_unknowns"] == g.n_nodes - wc.size + 1
