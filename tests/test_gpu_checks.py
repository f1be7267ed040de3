"""Closed-form self-checks on the device, with the bars of the reference's
own tests (tests/test_solvers.py:24-56): parallel plate and series stack to
1e-10, a capacitance matrix symmetric and conservative to 1e-9, the built-in
potential of twelve pn junctions to 1e-9 in at most 25 Newton steps; and the
isolated sphere converging to 4 pi eps a from above as the box grows."""

import numpy as np
import pytest

import fs_oracle as O
from fieldspice_amd.grid import RectilinearGrid

from fieldspice_amd.solvers import SolverConfig, checks
from fieldspice_amd.units import eps0, kB, q

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[gpu case] {msg}", flush=True)


@pytest.mark.parametrize("solver", ["hip_pcg", "hip_mixed"])
def test_parallel_plate(solver):
    _case(f"parallel plate, {solver}")
    r = checks.verify_parallel_plate(config=SolverConfig(linear_solver=solver))
    assert r["rel_err"] < 1e-10
    assert r["C"] == pytest.approx(r["C_ref"], rel=1e-10)
    # transverse variation is iteration error only: bounded by the 1e-10 tolerance
    assert r["uniform_dev"] < 1e-9
    rep = r["report"]
    assert rep["asymmetry_rel"] < 1e-9
    assert rep["row_sum_rel"] < 1e-9
    assert rep["linear"]["symmetric"] is True


def test_series_stack():
    _case("series dielectric stack")
    r = checks.verify_series_stack()
    assert r["rel_err"] < 1e-10
    assert r["v_interface"] == pytest.approx(r["v_interface_ref"], rel=1e-9)


def test_built_in_potential_sweep():
    _case("built-in potential, 12 junctions")
    r = checks.verify_built_in_potential_sweep()
    assert r["max_rel_err"] < 1e-9, r["max_rel_err"]
    assert r["max_iterations"] <= 25
    for c in r["cases"]:
        assert c["rel_err"] < 1e-9, c
        # the neutral-bulk difference the discrete equations solve exactly
        assert c["rel_err_exact"] < 1e-11, c


def test_built_in_potential_default_junction():
    _case("built-in potential, 1e17 cm^-3 symmetric junction")
    r = checks.verify_built_in_potential()
    assert r["rel_err"] < 1e-9
    assert r["uniform_dev"] < 1e-12
    # charge neutrality of the whole bar (unit cross-section), against the
    # dopant charge of one 45-Debye-length side
    LD = np.sqrt(11.7 * eps0 * kB * 300.0 / (q * q * 1e23))
    assert abs(r["net_charge"]) < 1e-6 * q * 1e23 * 45.0 * LD


def test_sphere_converges_from_above():
    _case("isolated sphere vs box padding")
    r = checks.sphere_capacitance_convergence(config=SolverConfig(tol=1e-12))
    C, err = np.array(r["C"]), np.array(r["rel_err"])
    assert np.all(np.diff(C) < 0), C               # the shield recedes: C falls
    assert np.all(err > 0), err                     # and stays above 4 pi eps a
    # a cube of half-width R lies outside the sphere of radius R, so the box
    # shields less than the concentric shell: C < 4 pi eps a / (1 - a/R), and
    # the de-shielded value approaches 4 pi eps a from below as R grows
    assert np.all(C < np.array(r["shell_pred"]))
    desh = np.array(r["deshielded"]) / r["C_ref"] - 1.0
    assert np.all(desh < 0) and np.all(np.diff(np.abs(desh)) < 0), desh
    assert abs(desh[-1]) < 0.04


@pytest.mark.parametrize("pad", [2.0, 3.0])
def test_sphere_octant_matches_oracle(pad):
    _case(f"sphere octant capacitance vs the oracle, pad {pad}")
    a = 1e-6
    r = checks.sphere_capacitance_convergence(pads=(pad,), config=SolverConfig(tol=1e-12))
    ax = checks._graded_axis(pad * a, a, a / 14, max(a, pad * a / 6.0), 1.25)
    g = RectilinearGrid(ax, ax, ax)
    X, Y, Z = g.node_coords()
    inside = np.flatnonzero(((X ** 2 + Y ** 2 + Z ** 2) <= a * a * (1 + 1e-12)).ravel())
    nid = g.node_ids()
    walls = np.unique(np.concatenate([nid[-1].ravel(), nid[:, -1].ravel(), nid[:, :, -1].ravel()]))
    Cr, _ = O.capacitance_matrix(g, np.full(g.shape_cells, eps0), [inside], wall_fixed=walls)
    assert r["n_nodes"][0] == g.n_nodes
    assert r["C"][0] == pytest.approx(8.0 * Cr[0, 0], rel=1e-9)
