"""Periodic walls on the device: the reference's wall-pair merge
(fieldspice/solvers/poisson.py:665-719,942-956) as a reduced grid operator
whose edges wrap around, against the oracle's merge (P^T L P, direct solve)."""

import numpy as np
import pytest

import fs_oracle as O
from fieldspice_amd.boundaries import BoundarySpec, Dirichlet, Periodic
from fieldspice_amd.grid import RectilinearGrid
from fieldspice_amd.solvers import PoissonSolver, SolverConfig
from fieldspice_amd.solvers.base import Terminal
from fieldspice_amd.solvers.poisson import NonlinearPoissonSolver, boltzmann_equilibrium_charge
from fieldspice_amd.units import eps0

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[gpu case] {msg}", flush=True)


def _torus(shape=(14, 11, 9), seed=6):
    g = RectilinearGrid(np.linspace(0, 1.3e-6, shape[0] + 1), np.linspace(0, 1.0e-6, shape[1] + 1),
                        np.linspace(0, 0.8e-6, shape[2] + 1))
    rng = np.random.default_rng(seed)
    return g, g.node_ids(), rng.uniform(2.0, 7.0, g.shape_cells) * eps0


def _wrap(nid, axis_len, xs, ys, zs):
    """Nodes of a box given by reduced x indices (taken mod axis_len), both
    copies of the merged wall plane included."""
    xs = sorted({i % axis_len for i in xs})
    full = xs + ([axis_len] if 0 in xs else [])
    return nid[full][:, ys][:, :, zs].ravel()


@pytest.mark.parametrize("solver", ["hip_pcg", "hip_mixed"])
@pytest.mark.parametrize("axes", [(0,), (0, 1)])
def test_periodic_capacitance_matches_oracle_merge(solver, axes):
    _case(f"periodic capacitance, axes {axes}, {solver}")
    g, nid, eps = _torus()
    Nx = g.shape_cells[0]
    walls = {}
    for a in axes:
        walls["xyz"[a] + "lo"] = Periodic()
        walls["xyz"[a] + "hi"] = Periodic()
    walls["zlo"] = Dirichlet(0.0)                       # a grounded shield
    bc = BoundarySpec(**walls)
    # one box straddles the periodic x wall
    terms = [Terminal("a", _wrap(nid, Nx, range(12, 17), range(2, 5), range(2, 5))),
             Terminal("b", _wrap(nid, Nx, range(5, 8), range(6, 9), range(3, 6))),
             Terminal("c", nid[9:11, 1:3, 6:8].ravel())]
    ps = PoissonSolver(g, eps, SolverConfig(linear_solver=solver, tol=1e-12))
    rep = {}
    C = ps.capacitance_matrix(terms, bc, report=rep)
    Cr, qsh = O.capacitance_matrix_merged(g, eps, [t.nodes for t in terms],
                                          O.periodic_sets(g, axes), wall_fixed=nid[:, :, 0].ravel())
    assert np.abs(C - Cr).max() <= 1e-9 * np.abs(Cr).max()
    assert np.abs(rep["shield_charge"] - qsh).max() <= 1e-9 * np.abs(Cr).max()


def test_periodic_solve_with_floating_electrode_matches_oracle_merge():
    _case("periodic potential with a driven box, a charged floating box and space charge")
    g, nid, eps = _torus(seed=9)
    Nx = g.shape_cells[0]
    bc = BoundarySpec(xlo=Periodic(), xhi=Periodic(), zhi=Dirichlet(0.2))
    drv = _wrap(nid, Nx, range(13, 16), range(3, 6), range(2, 4))
    flt = nid[6:9, 5:8, 3:5].ravel()
    rho = np.zeros(g.shape_nodes)
    rho[2:5, 1:3, 1:3] = 4e3
    Q = -3e-18
    ps = PoissonSolver(g, eps, SolverConfig(tol=1e-12))
    r = ps.solve([Terminal("drv", drv, voltage=0.8), Terminal("flt", flt, current=Q)], bc,
                 rho_node=rho)
    wall_idx, wall_val = bc.dirichlet_nodes(g)
    fixed, vals = O.unique_last(np.concatenate([wall_idx, drv]),
                                np.concatenate([wall_val, np.full(drv.size, 0.8)]))
    ref = O.poisson_floating(g, eps, fixed, vals, [flt], [Q],
                             b_free=rho.ravel() * g.node_volumes().ravel(),
                             merge=O.periodic_sets(g, (0,)))
    phi = r.fields["phi"][0].ravel()
    np.testing.assert_allclose(phi, ref, rtol=0, atol=1e-9 * np.abs(ref).max())
    lo, hi = O.periodic_pairs(g, 0)
    assert np.array_equal(phi[lo], phi[hi])
    assert r.terminals["flt"]["q"][0] == pytest.approx(Q, rel=1e-8)


def test_periodic_nonlinear_poisson_matches_oracle_newton():
    _case("periodic nonlinear Poisson (Boltzmann charge) vs the oracle's Newton")
    g = RectilinearGrid(np.linspace(0, 0.4e-6, 17), np.linspace(0, 0.3e-6, 13),
                        np.linspace(0, 0.2e-6, 9))
    nid = g.node_ids()
    eps = np.full(g.shape_cells, 11.7 * eps0)
    dop = np.full(g.shape_nodes, 1e22)
    dop[4:10, 3:9, :] = -1e22
    eq = boltzmann_equilibrium_charge(dop.ravel(), 1.5e16)
    contact = nid[:, :, 0].ravel()
    bc = BoundarySpec(xlo=Periodic(), xhi=Periodic(), ylo=Periodic(), yhi=Periodic())
    ns = NonlinearPoissonSolver(g, eps, eq.rho, eq.drho_dphi, SolverConfig(tol=1e-12))
    vc = float(eq.phi_guess[contact[0]])
    r = ns.solve([Terminal("c", contact, voltage=vc)], bc, phi0=eq.phi_guess)
    ref, _, _ = O.nonlinear_poisson(g, eps, eq.rho, eq.drho_dphi, fixed=contact,
                                    vals=np.full(contact.size, vc), phi0=eq.phi_guess,
                                    merge=O.periodic_sets(g, (0, 1)))
    phi = r.fields["phi"][0].ravel()
    assert np.abs(phi - ref).max() <= 1e-8 * np.abs(ref).max()
