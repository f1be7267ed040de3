"""GPU extraction callers beside the capacitance path: resistance matrix and
per-unit-length RLGC of a 2D cross-section, against closed forms, the
reference's own test values and the oracle."""

import numpy as np
import pytest

import fs_oracle as O
from fieldspice_amd import extraction as EX
from fieldspice_amd.grid import RectilinearGrid
from fieldspice_amd.solvers.base import Terminal
from fieldspice_amd.units import c0, eps0, mu0

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[gpu case] {msg}", flush=True)


def _plate_line():
    # the cross-section of tests/test_quasistatic.py:173-181 of the reference
    er, W, Hgap = 4.0, 200e-6, 50e-6
    g = RectilinearGrid(np.linspace(0, Hgap, 40), np.linspace(0, W, 60), None)
    eps = np.full(g.shape_cells, er * eps0)
    sig = np.zeros(g.shape_cells)
    nid = np.arange(g.n_nodes).reshape(g.shape_nodes)
    return g, eps, sig, [Terminal("sig", nid[0].ravel()), Terminal("gnd", nid[-1].ravel())], \
        er, W, Hgap


def test_rlgc_matches_the_parallel_plate_line():
    _case("rlgc_2d parallel-plate line vs closed forms")
    g, eps, sig, terms, er, W, Hgap = _plate_line()
    r = EX.rlgc_2d(g, eps, sig, terms, reference="gnd")
    # values and tolerances of the reference's tests/test_quasistatic.py:184-191
    assert r["C"][0, 0] == pytest.approx(er * eps0 * W / Hgap, rel=1e-9)
    assert r["L"][0, 0] == pytest.approx(mu0 * Hgap / W, rel=1e-9)
    assert r["Z0"][0] == pytest.approx(np.sqrt((mu0 * Hgap / W) / (er * eps0 * W / Hgap)),
                                       rel=1e-9)
    assert r["v_p"][0] == pytest.approx(c0 / np.sqrt(er), rel=1e-9)
    assert r["eps_eff"][0] == pytest.approx(er, rel=1e-9)
    assert r["reference"] == "gnd"
    assert np.all(r["G"] == 0.0) and np.all(r["R"] == 0.0)
    # tests/test_quasistatic.py:194-197: L C = mu eps I for a TEM line
    assert EX.check_lc_identity(r, eps_r=er) < 1e-9
    np.testing.assert_allclose(EX.characteristic_impedance(r), r["Z0"], rtol=1e-12)


def test_rlgc_inhomogeneous_and_lossy_cross_section_vs_oracle():
    _case("rlgc_2d microstrip-like, lossy substrate, 3 conductors vs oracle")
    g = RectilinearGrid(np.linspace(0, 40e-6, 41), np.linspace(0, 60e-6, 31), None)
    nid = np.arange(g.n_nodes).reshape(g.shape_nodes)
    eps = np.full(g.shape_cells, eps0)
    eps[:, :12, :] = 4.2 * eps0                  # substrate below y = 24 um
    sig = np.zeros(g.shape_cells)
    sig[:, :12, :] = 0.05                        # lossy substrate
    terms = [Terminal("a", nid[10:14, 12].ravel()), Terminal("b", nid[24:28, 12].ravel()),
             Terminal("gnd", nid[:, 0].ravel())]
    r = EX.rlgc_2d(g, eps, sig, terms)           # reference: the last conductor
    tn = [t.nodes for t in terms]
    keep = np.ix_([0, 1], [0, 1])
    Cr, _ = O.capacitance_matrix(g, eps, tn)
    Ca, _ = O.capacitance_matrix(g, np.full_like(eps, eps0), tn)
    Gr, _ = O.capacitance_matrix(g, np.maximum(sig, 1e-30), tn)
    np.testing.assert_allclose(r["C"], Cr[keep], rtol=1e-9, atol=1e-9 * np.abs(Cr).max())
    np.testing.assert_allclose(r["C_air"], Ca[keep], rtol=1e-9, atol=1e-9 * np.abs(Ca).max())
    np.testing.assert_allclose(r["G"], Gr[keep], rtol=1e-8, atol=1e-8 * np.abs(Gr).max())
    np.testing.assert_allclose(r["L"], mu0 * eps0 * np.linalg.inv(Ca[keep]), rtol=1e-9)
    assert np.all((1.0 < r["eps_eff"]) & (r["eps_eff"] < 4.2))
    # series resistance: every conducting cell is the conductor's area
    area = float(g.cell_volumes()[sig > 0].sum())
    np.testing.assert_allclose(np.diag(r["R"]), 1.0 / (0.05 * area), rtol=1e-12)
    z = EX.characteristic_impedance(r, freq=1e9)
    assert z.shape == (2,) and np.all(np.real(z) > 0)


def test_resistance_matrix_of_a_slab_is_the_pseudo_inverse_of_one_conductance():
    _case("resistance_matrix uniform slab vs R = d/(sigma A)")
    d, W, H, sigma = 2e-6, 5e-6, 4e-6, 1e4
    g = RectilinearGrid(np.linspace(0, d, 16), np.linspace(0, W, 6), np.linspace(0, H, 5))
    nid = g.node_ids()
    R = EX.resistance_matrix(g, np.full(g.shape_cells, sigma),
                             [Terminal("hi", nid[0].ravel()), Terminal("lo", nid[-1].ravel())])
    Rs = d / (sigma * W * H)
    # Y = G [[1, -1], [-1, 1]]  ->  pinv(Y) = Rs/4 [[1, -1], [-1, 1]]
    np.testing.assert_allclose(R, Rs / 4 * np.array([[1.0, -1.0], [-1.0, 1.0]]), rtol=1e-9)


def test_resistance_matrix_three_terminals_vs_oracle():
    _case("resistance_matrix 3 terminals, conductor with insulating gaps vs oracle")
    g = RectilinearGrid(np.linspace(0, 3e-6, 25), np.linspace(0, 2e-6, 17),
                        np.linspace(0, 1e-6, 9))
    rng = np.random.default_rng(11)
    sig = rng.uniform(1e3, 5e4, g.shape_cells)
    sig[10:14, 4:12, :] = 0.0                    # an insulating block
    nid = g.node_ids()
    terms = [Terminal("a", nid[0, :, :].ravel()), Terminal("b", nid[-1, :8, :].ravel()),
             Terminal("c", nid[-1, 10:, :].ravel())]
    R = EX.resistance_matrix(g, sig, terms)
    Y, _ = O.capacitance_matrix(g, sig, [t.nodes for t in terms])   # direct solves
    # the reference's pinv(Y) inverts Y's rounding residue along (1, 1, 1)
    # here (Y 1 ~ 1e-16 |Y|, above pinv's default cut-off): project it out
    P = np.eye(3) - 1.0 / 3.0
    Rr = np.linalg.pinv(P @ Y @ P, rcond=1e-8, hermitian=True)
    np.testing.assert_allclose(R, Rr, rtol=1e-8, atol=1e-8 * np.abs(Rr).max())
    # the Moore-Penrose identities: R Y R = R, and R Y the projector P
    np.testing.assert_allclose(R @ Y @ R, R, atol=1e-8 * np.abs(R).max())
    np.testing.assert_allclose(R @ Y, P, atol=1e-8)
    assert np.abs(R.sum(axis=1)).max() < 1e-9 * np.abs(R).max()
