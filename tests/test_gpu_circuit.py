"""GPU field region reduced onto its terminals, against the reference's
coupling tests (tests/test_quasistatic.py:243-330): the circuit around the
region is assembled by hand from the region's own stamps, since the netlist /
MNA machinery is not part of this package."""

import numpy as np
import pytest

from fieldspice_amd.circuit import FieldRegion
from fieldspice_amd.grid import RectilinearGrid
from fieldspice_amd.solvers.base import Terminal
from fieldspice_amd.units import eps0

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[gpu case] {msg}", flush=True)


def _plates(d=2e-6, W=5e-6, H=4e-6):
    g = RectilinearGrid(np.linspace(0, d, 16), np.linspace(0, W, 6), np.linspace(0, H, 5))
    nid = np.arange(g.n_nodes).reshape(g.shape_nodes)
    return g, [Terminal("a", nid[0].ravel()), Terminal("b", nid[-1].ravel())]


def test_reduced_admittance_reproduces_the_field_resistance():
    _case("field region: DC admittance = 1/R of the slab")
    d, W, H, sigma = 2e-6, 5e-6, 4e-6, 1e4
    g, terms = _plates(d, W, H)
    fr = FieldRegion("FR", g, np.full(g.shape_cells, 3.9 * eps0),
                     np.full(g.shape_cells, sigma), terms, circuit_nodes=["a", "b"])
    Y = fr.terminal_admittance(None)
    assert 1.0 / Y[0, 0] == pytest.approx(d / (sigma * W * H), rel=1e-9)
    assert fr.asymmetry < 1e-12


def test_reduced_admittance_reproduces_the_field_capacitance():
    _case("field region: Y dt = C of an insulating slab")
    d, W, H, er = 2e-6, 5e-6, 4e-6, 3.9
    g, terms = _plates(d, W, H)
    fr = FieldRegion("FR", g, np.full(g.shape_cells, er * eps0), np.zeros(g.shape_cells),
                     terms, circuit_nodes=["a", "b"])
    dt = 1e-12
    assert fr.terminal_admittance(dt)[0, 0] * dt == pytest.approx(er * eps0 * W * H / d,
                                                                 rel=1e-9)
    # sigma == 0: no DC conductance at all -- an open circuit, exactly
    assert np.abs(fr.terminal_admittance(None)).max() == 0.0


def test_coupled_resistive_divider_is_exact():
    _case("field region: resistive divider from the DC stamp")
    d, W, H, sigma = 2e-6, 5e-6, 4e-6, 1e4
    g, terms = _plates(d, W, H)
    Rf = d / (sigma * W * H)
    fr = FieldRegion("FR", g, np.full(g.shape_cells, 3.9 * eps0),
                     np.full(g.shape_cells, sigma), terms, circuit_nodes=["mid", "0"])
    nmap = {"mid": 0, "0": -1}
    G, I = np.zeros((1, 1)), np.zeros(1)
    fr.stamp_dc(G, I, None, nmap)
    G[0, 0] += 1.0 / Rf                 # R1 = Rf from a 1 V source into "mid"
    I[0] += 1.0 / Rf
    assert float(I[0] / G[0, 0]) == pytest.approx(0.5, abs=1e-9)


def test_coupled_rc_transient_matches_analytic():
    _case("field region: RC step response with backward Euler, first order in dt")
    d, W, H, er = 2e-6, 5e-6, 4e-6, 3.9
    g, terms = _plates(d, W, H)
    Cf = er * eps0 * W * H / d
    R = 1e6
    tau = R * Cf
    t0 = 2 * tau
    errs = []
    for spt in (25, 100):
        fr = FieldRegion("FR", g, np.full(g.shape_cells, er * eps0), np.zeros(g.shape_cells),
                         terms, circuit_nodes=["out", "0"])
        nmap = {"out": 0, "0": -1}
        dt = tau / spt
        x = np.zeros(1)
        ts, vs, its = [], [], []
        for k in range(1, 10 * spt + 1):
            t = k * dt
            vin = 1.0 if t >= t0 else 0.0
            G, I = np.array([[1.0 / R]]), np.array([vin / R])
            fr.stamp_tran(G, I, x, x, dt, nmap)
            x_new = I / G[0]
            fr.accept_timestep(x_new, x, dt, nmap)
            its.append(fr.last_iterations)
            x = x_new
            ts.append(t)
            vs.append(float(x[0]))
        ts, vs = np.array(ts), np.array(vs)
        m = ts >= t0
        errs.append(np.abs(vs[m] - (1.0 - np.exp(-(ts[m] - t0) / tau))).max())
    assert errs[-1] < 2.4e-2
    assert 3.0 < errs[0] / errs[1] < 5.0      # first order: 4x the steps, 4x smaller error
    # a settled step's back-substitution started from the previous field
    # needs fewer iterations than the same solve started from zero
    entry = fr._reduce(dt)
    f = fr.L_eps.apply(fr.phi) / dt
    v = np.array([float(x[0]), 0.0])            # electrode voltages: "out", ground
    cold = fr._interior_solve(entry, f, v)
    n_cold = entry["ls"].info["iterations"]
    warm = fr._interior_solve(entry, f, v, guess=fr.phi)
    n_warm = entry["ls"].info["iterations"]
    print(f"[gpu case] back-substitution iterations: cold {n_cold}, warm {n_warm}, "
          f"in the run {its[:2]}...{its[-2:]}", flush=True)
    assert n_warm < n_cold
    np.testing.assert_allclose(warm, cold, rtol=0, atol=1e-9 * np.abs(cold).max())


def test_ac_stamp_of_a_lossy_slab_is_the_parallel_rc():
    _case("field region: AC stamp of a uniform lossy slab = (sigma + j w eps) A/d")
    d, W, H, er, sigma = 2e-6, 5e-6, 4e-6, 3.9, 1e-1
    g, terms = _plates(d, W, H)
    fr = FieldRegion("FR", g, np.full(g.shape_cells, er * eps0), np.full(g.shape_cells, sigma),
                     terms, circuit_nodes=["in", "0"])
    w = 2e9 * np.pi
    y = (sigma + 1j * w * er * eps0) * W * H / d
    Ym, J = np.zeros((1, 1), complex), np.zeros(1, complex)
    fr.stamp_ac(Ym, J, None, w, {"in": 0, "0": -1})     # ground dropped
    assert Ym[0, 0] == pytest.approx(y, rel=1e-9)
    assert not np.any(J)
    np.testing.assert_allclose(fr.ac_admittance(-w), np.conj(fr.ac_admittance(w)), rtol=0)
    np.testing.assert_allclose(fr.ac_admittance(0.0), fr.terminal_admittance(None), rtol=0)
    assert fr.ac_info["iterations"] <= 30


@pytest.mark.parametrize("sig_scale", [1e-4, 1.0])
def test_ac_admittance_of_an_inhomogeneous_region_vs_oracle(sig_scale):
    _case(f"field region: AC admittance, 3 electrodes, random eps/sigma x{sig_scale} vs oracle")
    import fs_oracle as O
    g = RectilinearGrid(np.linspace(0, 3e-6, 25), np.linspace(0, 2e-6, 17),
                        np.linspace(0, 1e-6, 9))
    rng = np.random.default_rng(21)
    eps = rng.uniform(1.0, 8.0, g.shape_cells) * eps0
    sig = sig_scale * rng.uniform(1e3, 5e4, g.shape_cells)
    sig[10:14, 4:12, :] = 0.0                    # an insulating block
    nid = g.node_ids()
    terms = [Terminal("a", nid[0].ravel()), Terminal("b", nid[-1, :8].ravel()),
             Terminal("c", nid[-1, 10:].ravel())]
    fr = FieldRegion("FR", g, eps, sig, terms)
    w = 2e9 * np.pi
    Y = fr.ac_admittance(w)
    Yr = O.field_region_ac(g, eps, sig, [t.nodes for t in terms], w)
    np.testing.assert_allclose(Y, Yr, rtol=1e-8, atol=1e-8 * np.abs(Yr).max())
    assert fr.ac_info["iterations"] <= 40
    # the MNA view: stamped into a 2-node system with "c" grounded
    fr2 = FieldRegion("FR2", g, eps, sig, terms, circuit_nodes=["n1", "n2", "0"])
    Ym = np.zeros((2, 2), complex)
    fr2.stamp_ac(Ym, np.zeros(2, complex), None, w, {"n1": 0, "n2": 1, "0": -1})
    np.testing.assert_allclose(Ym, Yr[:2, :2], rtol=1e-8, atol=1e-8 * np.abs(Yr).max())
