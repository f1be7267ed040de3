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
        ts, vs = [], []
        for k in range(1, 10 * spt + 1):
            t = k * dt
            vin = 1.0 if t >= t0 else 0.0
            G, I = np.array([[1.0 / R]]), np.array([vin / R])
            fr.stamp_tran(G, I, x, x, dt, nmap)
            x_new = I / G[0]
            fr.accept_timestep(x_new, x, dt, nmap)
            x = x_new
            ts.append(t)
            vs.append(float(x[0]))
        ts, vs = np.array(ts), np.array(vs)
        m = ts >= t0
        errs.append(np.abs(vs[m] - (1.0 - np.exp(-(ts[m] - t0) / tau))).max())
    assert errs[-1] < 2.4e-2
    assert 3.0 < errs[0] / errs[1] < 5.0      # first order: 4x the steps, 4x smaller error
    with pytest.raises(NotImplementedError):
        fr.stamp_ac(np.zeros((1, 1), complex), np.zeros(1, complex), None, 1e9, nmap)
