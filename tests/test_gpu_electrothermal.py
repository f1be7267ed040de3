"""GPU electro-thermal coupling against the lumped analytic fixed points of the
reference's tests/test_thermal.py:237-349 (same geometry, values and
tolerances)."""

import warnings

import numpy as np
import pytest

from fieldspice_amd.grid import RectilinearGrid
from fieldspice_amd.physics import PhysicsOptions
from fieldspice_amd.solvers.base import Terminal
from fieldspice_amd.solvers.electrothermal import (ElectroThermalSolver, ThermalRunaway,
                                                   runaway_threshold)

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[gpu case] {msg}", flush=True)


def _lump():
    """A bar that is isothermal by construction (kappa 5e5), so the lumped
    fixed point applies exactly."""
    Lc, Wc, Hc = 100e-6, 20e-6, 20e-6
    sigma0, alpha, h = 1e6, 3.93e-3, 2e4
    g = RectilinearGrid(np.linspace(0, Lc, 21), np.linspace(0, Wc, 5), np.linspace(0, Hc, 5))
    sig = np.full(g.shape_cells, sigma0)
    kap = np.full(g.shape_cells, 5e5)
    A_wall = 2 * (Lc * Wc + Lc * Hc + Wc * Hc)
    R_th = 1.0 / (h * A_wall)
    R0 = Lc / (sigma0 * Wc * Hc)
    conv = {w: (h, 300.0) for w in ("xlo", "xhi", "ylo", "yhi", "zlo", "zhi")}
    nid = np.arange(g.n_nodes).reshape(g.shape_nodes)
    return g, sig, kap, alpha, R_th, R0, conv, nid


def _solve(term_a, g, sig, kap, alpha, conv, nid, T_limit=1e9, damping=0.6):
    tb = Terminal("b", nid[-1].ravel(), voltage=0.0)
    o = PhysicsOptions(self_heating=True, max_coupling_iterations=3000, coupling_tolerance=1e-10)
    s = ElectroThermalSolver(g, sig, kap, alpha, o)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return s.solve([term_a, tb], convection=conv, damping=damping, T_limit=T_limit)


@pytest.mark.parametrize("V", [0.1, 0.5, 1.0, 2.0, 5.0])
def test_voltage_driven_matches_the_analytic_fixed_point(V):
    _case(f"electro-thermal voltage drive {V} V")
    g, sig, kap, alpha, R_th, R0, conv, nid = _lump()
    r = _solve(Terminal("a", nid[0].ravel(), voltage=V), g, sig, kap, alpha, conv, nid)
    B = R_th * V * V / R0          # alpha dT^2 + dT - B = 0
    ana = (-1 + np.sqrt(1 + 4 * alpha * B)) / (2 * alpha)
    assert float(r.scalars["T_rise"][0]) == pytest.approx(ana, rel=1e-5)


@pytest.mark.parametrize("frac", [0.2, 0.5, 0.9])
def test_current_driven_matches_the_analytic_fixed_point(frac):
    _case(f"electro-thermal current drive {frac} of the runaway threshold")
    g, sig, kap, alpha, R_th, R0, conv, nid = _lump()
    I = frac * runaway_threshold(R_th, R0, alpha)
    r = _solve(Terminal("a", nid[0].ravel(), current=I), g, sig, kap, alpha, conv, nid)
    A = R_th * I * I * R0          # dT = A/(1 - A alpha)
    assert float(r.scalars["T_rise"][0]) == pytest.approx(A / (1 - A * alpha), rel=1e-4)
    assert float(r.terminals["a"]["i"][0]) == pytest.approx(I, rel=1e-8)


@pytest.mark.parametrize("frac", [1.01, 1.1, 1.5, 3.0])
def test_current_drive_above_threshold_runs_away(frac):
    _case(f"electro-thermal runaway at {frac} of the threshold")
    g, sig, kap, alpha, R_th, R0, conv, nid = _lump()
    I = frac * runaway_threshold(R_th, R0, alpha)
    with pytest.raises(ThermalRunaway):
        _solve(Terminal("a", nid[0].ravel(), current=I), g, sig, kap, alpha, conv, nid,
               T_limit=5000.0)


def test_voltage_drive_is_never_declared_a_runaway():
    _case("electro-thermal voltage drive overshoot is not a runaway")
    g, sig, kap, alpha, R_th, R0, conv, nid = _lump()
    r = _solve(Terminal("a", nid[0].ravel(), voltage=1.0), g, sig, kap, alpha, conv, nid,
               T_limit=5000.0)
    assert float(r.scalars["T_rise"][0]) == pytest.approx(2280.93, rel=1e-3)
    assert r.scalars["history"].max() > 5000.0


def test_zero_tcr_makes_the_coupling_one_pass_and_records_assumptions():
    _case("electro-thermal without feedback, and the assumption record")
    g, sig, kap, alpha, R_th, R0, conv, nid = _lump()
    r = _solve(Terminal("a", nid[0].ravel(), voltage=0.5), g, sig, kap, 0.0, conv, nid,
               damping=1.0)
    assert float(r.scalars["T_rise"][0]) == pytest.approx(R_th * 0.5 ** 2 / R0, rel=1e-4)
    assert r.meta["iterations"] <= 3
    assert "A6" not in r.meta["assumptions"] and "A1a" in r.meta["assumptions"]
    assert "self_heating" in r.meta["physics"]


def test_floating_terminal_is_left_unconstrained():
    """A floating terminal constrains nothing in the conduction problem (the
    reference assembles only voltage and current terminals,
    fieldspice/solvers/electrothermal.py:143-173): the fields equal those of
    the run without it, and it reports the dual-volume weighted potential of
    its nodes and no net current."""
    _case("electro-thermal run with a floating probe terminal")
    g, sig, kap, alpha, R_th, R0, conv, nid = _lump()
    ta = Terminal("a", nid[0].ravel(), voltage=0.5)
    probe = Terminal("probe", nid[10, 1:3, 1:3].ravel(), voltage=None)
    assert probe.driven == "float"
    r0 = _solve(ta, g, sig, kap, alpha, conv, nid)
    tb = Terminal("b", nid[-1].ravel(), voltage=0.0)
    o = PhysicsOptions(self_heating=True, max_coupling_iterations=3000, coupling_tolerance=1e-10)
    s = ElectroThermalSolver(g, sig, kap, alpha, o)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        r1 = s.solve([ta, tb, probe], convection=conv, damping=0.6, T_limit=1e9)
    assert np.array_equal(r1.fields["phi"], r0.fields["phi"])
    assert np.array_equal(r1.fields["T"], r0.fields["T"])
    phi = r1.fields["phi"][0]
    w = g.node_volumes().ravel()[probe.nodes]
    assert r1.v("probe")[0] == pytest.approx(float(w @ phi[probe.nodes]) / w.sum(), rel=1e-14)
    assert 0.0 < r1.v("probe")[0] < 0.5
    assert abs(r1.i("probe")[0]) <= 1e-8 * abs(r1.i("a")[0])
