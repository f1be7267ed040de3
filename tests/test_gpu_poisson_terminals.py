"""GPU Poisson solve with floating and charge-driven electrodes against the
oracle's node merge (``fieldspice/solvers/poisson.py:1039-1063``) and the
series-capacitor closed form."""

import numpy as np
import pytest

import fs_oracle as O
from fieldspice_amd.boundaries import BoundarySpec, Dirichlet
from fieldspice_amd.grid import RectilinearGrid
from fieldspice_amd.solvers import PoissonSolver, SolverConfig
from fieldspice_amd.solvers.base import Terminal
from fieldspice_amd.units import eps0

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[gpu case] {msg}", flush=True)


def test_floating_plate_at_the_series_capacitor_potential():
    _case("poisson: floating / charged full-plane plate between driven plates")
    d, W, H, er = 3e-6, 2e-6, 1e-6, 3.0
    xs = np.linspace(0, d, 25)
    g = RectilinearGrid(xs, np.linspace(0, W, 9), np.linspace(0, H, 5))
    nid = g.node_ids()
    ps = PoissonSolver(g, np.full(g.shape_cells, er * eps0))
    ia = 9
    C1 = er * eps0 * W * H / xs[ia]
    C2 = er * eps0 * W * H / (d - xs[ia])
    for Q in (None, 3e-17):
        plate = Terminal("f", nid[ia].ravel(), current=Q)
        r = ps.solve([Terminal("hi", nid[0].ravel(), voltage=1.0), plate,
                      Terminal("lo", nid[-1].ravel(), voltage=0.0)])
        q = 0.0 if Q is None else Q
        assert r.terminals["f"]["v"][0] == pytest.approx((q + C1) / (C1 + C2), rel=1e-9)
        assert r.terminals["f"]["q"][0] == pytest.approx(q, abs=1e-9 * C1)
        phi = r.fields["phi"][0]
        assert np.ptp(phi[ia]) < 1e-12            # the plate is one equipotential
        # the driven plates' charges balance the floating one's
        assert (r.terminals["hi"]["q"][0] + r.terminals["lo"]["q"][0]
                == pytest.approx(-q, abs=1e-9 * C1))
        assert r.meta["n_unknowns"] == g.n_nodes - nid[ia].size + 1


@pytest.mark.parametrize("solver", ["hip_pcg", "hip_mixed"])
def test_floating_and_charged_boxes_vs_oracle_merge(solver):
    _case(f"poisson: floating + charged boxes, wall, space charge vs oracle ({solver})")
    g = RectilinearGrid(np.linspace(0, 4e-6, 33), np.linspace(0, 3e-6, 25),
                        np.linspace(0, 2e-6, 17))
    rng = np.random.default_rng(8)
    eps = rng.uniform(2.0, 9.0, g.shape_cells) * eps0
    nid = g.node_ids()
    drv = nid[4:8, 4:8, 4:8].ravel()
    flt = nid[14:20, 10:16, 6:10].ravel()
    chg = nid[24:28, 16:20, 8:12].ravel()
    rho = np.zeros(g.shape_nodes)
    rho[18:22, 4:8, 10:14] = 2e3                      # C/m^3
    bc = BoundarySpec(xlo=Dirichlet(0.0))
    Qc = -2e-17
    ps = PoissonSolver(g, eps, SolverConfig(linear_solver=solver, tol=1e-12))
    r = ps.solve([Terminal("drv", drv, voltage=0.7), Terminal("flt", flt),
                  Terminal("chg", chg, current=Qc)], bc, rho_node=rho)
    wall = nid[0].ravel()
    fixed = np.concatenate([wall, drv])
    vals = np.concatenate([np.zeros(wall.size), np.full(drv.size, 0.7)])
    b_free = rho.ravel() * g.node_volumes().ravel()
    ref = O.poisson_floating(g, eps, fixed, vals, [flt, chg], [0.0, Qc], b_free=b_free)
    phi = r.fields["phi"][0].ravel()
    np.testing.assert_allclose(phi, ref, rtol=0, atol=1e-9 * np.abs(ref).max())
    assert r.terminals["flt"]["q"][0] == pytest.approx(0.0, abs=1e-9 * abs(Qc))
    assert r.terminals["chg"]["q"][0] == pytest.approx(Qc, rel=1e-8)
    assert r.terminals["flt"]["v"][0] == pytest.approx(ref[flt[0]], rel=1e-9)


def test_floating_terminal_limits_are_explicit():
    _case("poisson: floating terminal with no reference / touching a wall")
    g = RectilinearGrid(np.linspace(0, 1e-6, 9), np.linspace(0, 1e-6, 9), np.linspace(0, 1e-6, 9))
    nid = g.node_ids()
    ps = PoissonSolver(g, np.full(g.shape_cells, eps0))
    with pytest.raises(NotImplementedError, match="no Dirichlet node"):
        ps.solve([Terminal("f", nid[3:5, 3:5, 3:5].ravel(), current=1e-18)])
    with pytest.raises(NotImplementedError, match="touches a Dirichlet wall"):
        ps.solve([Terminal("f", nid[0:2, 3:5, 3:5].ravel())], BoundarySpec(xlo=Dirichlet(0.0)))
