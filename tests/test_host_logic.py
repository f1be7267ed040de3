"""Host-side logic of the GPU solvers (no GPU): validation, Dirichlet groups,
source bases, material maps, boundary data."""

import numpy as np
import pytest

import fs_oracle as O
from fieldspice_amd.boundaries import BoundarySpec, Dirichlet, Neumann, wall_dual_areas, wall_nodes
from fieldspice_amd.grid import RectilinearGrid
from fieldspice_amd.materials import MaterialMap, get
from fieldspice_amd.operators import cell_to_node, edge_weights_from_property
from fieldspice_amd.solvers.base import SolverConfig, Terminal, check_terminals
from fieldspice_amd.solvers.transient import DirichletGroups, source_basis
from fieldspice_amd.units import eps0
from fieldspice_amd import workloads


def _g():
    return RectilinearGrid.uniform([(0, 1.0), (0, 2.0), (0, 3.0)], [4, 5, 6])


def test_terminal_validation_mirrors_reference_messages():
    g = _g()
    nid = g.node_ids()
    a = Terminal("a", nid[0].ravel())
    with pytest.raises(ValueError, match="bare Terminal"):
        check_terminals(a, g)
    with pytest.raises(ValueError, match="duplicate"):
        check_terminals([a, Terminal("a", nid[1].ravel())], g)
    with pytest.raises(ValueError, match="share node"):
        check_terminals([a, Terminal("b", nid[0, 0].ravel())], g)
    with pytest.raises(ValueError, match="outside"):
        check_terminals([Terminal("c", [g.n_nodes])], g)
    with pytest.raises(ValueError, match="voltage or current"):
        Terminal("d", [0], voltage=1.0, current=1.0)
    with pytest.raises(ValueError, match="linear_solver"):
        SolverConfig(linear_solver="bogus")


def test_wall_nodes_areas_and_precedence():
    g = _g()
    assert wall_dual_areas(g, "xlo").sum() == pytest.approx(2.0 * 3.0, rel=1e-12)
    assert np.all(np.diff(wall_nodes(g, "ylo")) > 0)
    bc = BoundarySpec(xlo=Dirichlet(0.0), zhi=Dirichlet(1.0))
    idx, val = bc.dirichlet_nodes(g)
    corner = g.node_index(0, 0, 6)
    assert val[np.searchsorted(idx, corner)] == 1.0   # later wall wins
    with pytest.raises(ValueError):
        BoundarySpec(xlo=1.0)


def test_dirichlet_groups_reproduce_last_wins_dedup():
    g = _g()
    nid = g.node_ids()
    t1 = Terminal("t1", nid[0, :2, :2].ravel(), voltage=lambda t: 2.0 * t)
    t2 = Terminal("t2", nid[3, 3, 3].ravel(), voltage=0.5)
    tf = Terminal("tf", nid[2, 2, 2].ravel())   # floating: not fixed
    bc = BoundarySpec(xlo=Dirichlet(-1.0))
    for walls_last in (True, False):
        G = DirichletGroups(g, [t1, t2, tf], bc, walls_last=walls_last)
        idx, vals = G.node_values(0.25)
        parts_i = [t1.nodes, t2.nodes]
        parts_v = [np.full(t1.nodes.size, 0.5), np.full(1, 0.5)]
        w_idx, w_val = bc.dirichlet_nodes(g)
        if walls_last:
            ri, rv = O.unique_last(np.concatenate(parts_i + [w_idx]), np.concatenate(parts_v + [w_val]))
        else:
            ri, rv = O.unique_last(np.concatenate([w_idx] + parts_i), np.concatenate([w_val] + parts_v))
        assert np.array_equal(idx, ri) and np.array_equal(vals, rv)


def test_source_basis_reconstructs_sequences():
    rng = np.random.default_rng(0)
    a, b = rng.standard_normal(50), rng.standard_normal(50)
    seq = [np.zeros(50), a, 2 * a, 0 * a, b, a - 0.0, 3 * b]
    basis, scales = source_basis(seq)
    assert basis.shape[0] == 2
    for v, s in zip(seq, scales):
        assert np.allclose(s @ basis, v)
    with pytest.raises(NotImplementedError):
        source_basis([rng.standard_normal(5) for _ in range(10)], max_basis=8)


def test_material_map_absolute_units_and_mixing():
    g = RectilinearGrid.uniform([(0, 1.0)], [2])
    mm = MaterialMap(g, background="sio2")
    assert mm.eps().min() == pytest.approx(3.9 * eps0)
    assert get("cu").sigma == pytest.approx(5.8e7)
    g3 = RectilinearGrid.uniform([(0, 1.0), (0, 1.0), (0, 1.0)], [2, 1, 1])
    mm = MaterialMap(g3, background="vacuum")
    fill = np.zeros(g3.shape_cells)
    fill[0] = 1.0
    fill[1] = 0.5
    mm.assign(fill, "si")
    assert mm.eps_r()[1, 0, 0] == pytest.approx(0.5 * (1.0 + 11.7))
    with pytest.raises(ValueError):
        mm.assign(np.zeros((2, 1)), "si")


def test_cell_to_node_and_edge_weights_match_oracle():
    g = RectilinearGrid(np.cumsum([0, 1, 2, 1.5, 3]), np.cumsum([0, 2, 1]), np.cumsum([0, 1, 1, 4]))
    rng = np.random.default_rng(1)
    p = rng.uniform(1, 5, g.shape_cells)
    assert np.allclose(cell_to_node(g, p), O.cell_to_node(g, p), rtol=1e-15)
    pe = O.cell_to_edge(g, p)
    assert np.allclose(edge_weights_from_property(g, pe), O.edge_weights(g, pe), rtol=1e-15)


def test_c2_geometry_generator():
    ex = workloads.c2_geometry(seed=0)
    g = ex.grid
    assert g.shape_nodes == (192, 192, 96)
    assert np.allclose(g.hx, 50e-9) and np.allclose(g.hy, 50e-9)
    assert g.hz[0] == pytest.approx(20e-9) and g.hz[-1] == pytest.approx(400e-9)
    assert len(ex.terminals) == 8
    assert set(np.unique(ex.eps_cell / eps0).round(6)) <= {2.7, 3.9, 7.5}
    for (i0, i1, j0, j1, k0, k1) in ex.boxes:
        for s in (i1 - i0, j1 - j0, k1 - k0):
            assert 4 <= s <= 40
    check_terminals(ex.terminals, g)   # disjoint node sets


def test_extraction_argument_checks_and_closed_forms():
    from fieldspice_amd import extraction as EX
    from fieldspice_amd.units import mu0
    # rlgc_2d refuses before any solve (tests/test_quasistatic.py:200-215 of the reference)
    g2 = RectilinearGrid(np.linspace(0, 5e-5, 6), np.linspace(0, 2e-4, 7), None)
    nid = np.arange(g2.n_nodes).reshape(g2.shape_nodes)
    terms = [Terminal("sig", nid[0].ravel()), Terminal("gnd", nid[-1].ravel())]
    eps = np.full(g2.shape_cells, eps0)
    sig = np.zeros(g2.shape_cells)
    with pytest.raises(ValueError, match="two conductors"):
        EX.rlgc_2d(g2, eps, sig, terms[:1])
    with pytest.raises(ValueError, match="reference"):
        EX.rlgc_2d(g2, eps, sig, terms, reference="nope")
    g3 = RectilinearGrid.uniform([(0, 1e-6), (0, 1e-6), (0, 1e-6)], [4, 4, 4])
    n3 = g3.node_ids()
    with pytest.raises(ValueError, match="2D cross-section"):
        EX.rlgc_2d(g3, np.full(g3.shape_cells, eps0), np.zeros(g3.shape_cells),
                   [Terminal("a", n3[0].ravel()), Terminal("b", n3[-1].ravel())])
    with pytest.raises(ValueError, match="at least two terminals"):
        EX.resistance_matrix(g3, np.ones(g3.shape_cells), [Terminal("a", n3[0].ravel())])
    with pytest.raises(ValueError, match="non-negative"):
        EX.resistance_matrix(g3, -np.ones(g3.shape_cells),
                             [Terminal("a", n3[0].ravel()), Terminal("b", n3[-1].ravel())])
    # skin depth: sqrt(2/(omega mu sigma)); copper at 1 GHz is 2.09 um
    assert EX.skin_depth(5.8e7, 1e9) == pytest.approx(2.0898e-6, rel=1e-4)
    assert EX.skin_depth(5.8e7, 1e9) == pytest.approx(
        np.sqrt(1.0 / (np.pi * 1e9 * mu0 * 5.8e7)), rel=1e-14)
    assert EX.skin_depth(5.8e7, 0.0) == float("inf")
    # Z0 of an RLGC set: lossless sqrt(L/C), lossy sqrt((R + jwL)/(G + jwC))
    r = {"R": np.diag([2.0]), "L": np.diag([3e-7]), "G": np.diag([1e-3]), "C": np.diag([1.2e-10])}
    assert EX.characteristic_impedance(r)[0] == pytest.approx(np.sqrt(3e-7 / 1.2e-10))
    w = 2 * np.pi * 1e8
    assert EX.characteristic_impedance(r, 1e8)[0] == pytest.approx(
        np.sqrt((2.0 + 1j * w * 3e-7) / (1e-3 + 1j * w * 1.2e-10)))
    # L C = mu eps I exactly for these
    er = 2.5
    C = np.array([[2.0, -0.5], [-0.5, 1.5]]) * 1e-10
    lc = {"L": mu0 * er * eps0 * np.linalg.inv(C), "C": C}
    assert EX.check_lc_identity(lc, eps_r=er) < 1e-14
    lc["L"] = lc["L"] * (1 + 1e-3)
    assert EX.check_lc_identity(lc, eps_r=er) == pytest.approx(1e-3, rel=1e-6)


def test_nonlinear_poisson_argument_checks_and_edge_gradient():
    from fieldspice_amd.solvers.poisson import (NonlinearPoissonSolver, boltzmann_equilibrium_charge,
                                                edge_gradient)
    g = _g()
    eps = np.full(g.shape_cells, 11.7 * eps0)
    f = lambda p: p  # noqa: E731
    with pytest.raises(ValueError, match="callables"):
        NonlinearPoissonSolver(g, eps, 1.0, f)
    with pytest.raises(ValueError, match="step_limit"):
        NonlinearPoissonSolver(g, eps, f, f, step_limit=0.0)
    with pytest.raises(ValueError, match="armijo"):
        NonlinearPoissonSolver(g, eps, f, f, armijo=0.5)
    with pytest.raises(ValueError, match="v_scale"):
        NonlinearPoissonSolver(g, eps, f, f, v_scale=-1.0)
    with pytest.raises(ValueError, match="ni"):
        boltzmann_equilibrium_charge(np.zeros(4), 0.0)
    # the charge model of the reference's doctest (poisson.py:1589-1596)
    ec = boltzmann_equilibrium_charge(np.array([-1e23, 1e23]), 9.65e15)
    assert np.all(np.abs(ec.rho(ec.phi_guess)) < 1e-9)
    rho, drho, guess, Vt = O.boltzmann_charge(np.array([-1e23, 3e21, 1e23]), 9.65e15)
    ec = boltzmann_equilibrium_charge(np.array([-1e23, 3e21, 1e23]), 9.65e15)
    p = np.array([-0.3, 0.1, 0.45])
    np.testing.assert_array_equal(ec.rho(p), rho(p))
    np.testing.assert_array_equal(ec.drho_dphi(p), drho(p))
    # drho is the derivative of rho (central difference)
    h = 1e-7
    np.testing.assert_allclose((ec.rho(p + h) - ec.rho(p - h)) / (2 * h), ec.drho_dphi(p),
                               rtol=1e-6)
    # G phi, head minus tail per edge family, equals the incidence matrix
    phi = np.random.default_rng(0).standard_normal(g.n_nodes)
    np.testing.assert_allclose(edge_gradient(g, phi), O.grad_node_edge(g) @ phi, rtol=0,
                               atol=1e-15)


def test_electrothermal_host_checks():
    from fieldspice_amd.physics import PhysicsOptions
    from fieldspice_amd.solvers.electrothermal import ElectroThermalSolver, runaway_threshold
    # tests/test_thermal.py:329-340 of the reference
    assert runaway_threshold(1e3, 1.0, 1e-3) == pytest.approx(1.0)
    assert runaway_threshold(1e3, 1.0, 0.0) == float("inf")
    assert runaway_threshold(1e3, 1.0, -1e-3) == float("inf")
    g = _g()
    sig, kap = np.ones(g.shape_cells), np.ones(g.shape_cells)
    with pytest.raises(ValueError, match="self_heating"):
        ElectroThermalSolver(g, sig, kap, 3.93e-3, PhysicsOptions())
    with pytest.raises(ValueError, match="tcr_cell"):
        ElectroThermalSolver(g, sig, kap, np.zeros(3))
    with pytest.raises(ValueError, match="coupling_tolerance"):
        PhysicsOptions(coupling_tolerance=0.0)
    o = PhysicsOptions(self_heating=True)
    assert o.sigma_varies_with_T and "A6" in o.relaxed_assumptions()
    assert o.remaining_assumptions(("A1a", "A6")) == ("A1a",)


def test_field_region_host_checks_and_stamps():
    import scipy.sparse as sp
    from fieldspice_amd.circuit import FieldRegion, stamp_g, stamp_i
    g = _g()
    nid = g.node_ids()
    eps = np.full(g.shape_cells, 3.9 * eps0)
    a, b = Terminal("a", nid[0].ravel()), Terminal("b", nid[-1].ravel())
    with pytest.raises(ValueError, match="at least one terminal"):
        FieldRegion("FR", g, eps, np.zeros(g.shape_cells), [])
    with pytest.raises(ValueError, match="circuit_nodes"):
        FieldRegion("FR", g, eps, np.zeros(g.shape_cells), [a, b], circuit_nodes=["x"])
    with pytest.raises(ValueError, match="ABSOLUTE"):
        FieldRegion("FR", g, -eps, np.zeros(g.shape_cells), [a, b])
    with pytest.raises(ValueError, match="overlap"):
        FieldRegion("FR", g, eps, np.zeros(g.shape_cells), [a, Terminal("c", nid[0, 0])])
    G = np.zeros((2, 2))
    stamp_g(G, 0, 1, 2.0)
    stamp_g(G, -1, 1, 5.0)             # ground dropped
    L = sp.dok_matrix((2, 2))
    stamp_g(L, 1, 1, 3.0)
    I = np.zeros(2)
    stamp_i(I, 1, 4.0)
    stamp_i(I, -1, 9.0)
    assert G[0, 1] == 2.0 and G.sum() == 2.0 and L[1, 1] == 3.0 and I.tolist() == [0.0, 4.0]


def test_presb_fgmres_solves_complex_symmetric_systems():
    """The AC stamp's Krylov solver on SciPy operators: the block system
    ``(B + jC) u = f + jg`` against a dense complex solve, for a conductor
    (B dominant), an insulator (B = 0) and a mixed case, several right-hand
    sides in lockstep."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    from fieldspice_amd.circuit import presb_fgmres
    g = RectilinearGrid(np.linspace(0, 1, 9), np.linspace(0, 1, 7), np.linspace(0, 1, 5))
    rng = np.random.default_rng(5)
    nid = g.node_ids()
    fixed = np.concatenate([nid[0].ravel(), nid[-1].ravel()])
    inner = np.setdiff1d(np.arange(g.n_nodes), fixed)
    Le = O.nodal_laplacian(g, O.cell_to_edge(g, rng.uniform(1, 5, g.shape_cells)))
    for sig_scale, w in ((1e3, 1.0), (0.0, 1.0), (1.0, 1.0), (1.0, 1e-6)):
        sig = sig_scale * rng.uniform(0, 2, g.shape_cells) * (rng.random(g.shape_cells) < 0.6)
        Ls = O.nodal_laplacian(g, O.cell_to_edge(g, sig))
        B = Ls[inner][:, inner].tocsc()
        Cm = (w * Le)[inner][:, inner].tocsc()
        lu = spl.splu((B + Cm).tocsc())
        f = rng.standard_normal((3, inner.size))
        gg = rng.standard_normal((3, inner.size))
        x, y, info = presb_fgmres(lambda X: (B @ X.T).T, lambda X: (Cm @ X.T).T,
                                  lambda X: lu.solve(np.ascontiguousarray(X.T)).T, f, gg,
                                  tol=1e-11)
        K = (B + 1j * Cm).toarray()
        u = np.linalg.solve(K, (f + 1j * gg).T).T
        np.testing.assert_allclose(x + 1j * y, u, rtol=0, atol=1e-9 * np.abs(u).max())
        assert info["iterations"] <= 30 and info["residual"] <= 1e-11
    # a zero right-hand side column stays zero; the others still converge
    f[1] = 0.0
    gg[1] = 0.0
    x, y, _ = presb_fgmres(lambda X: (B @ X.T).T, lambda X: (Cm @ X.T).T,
                           lambda X: lu.solve(np.ascontiguousarray(X.T)).T, f, gg)
    assert not np.any(x[1]) and not np.any(y[1])
    with pytest.raises(RuntimeError, match="did not reach"):
        presb_fgmres(lambda X: (B @ X.T).T, lambda X: (Cm @ X.T).T, lambda X: X, f, gg,
                     restart=2, max_restarts=1)


def test_bench_refuses_more_than_one_gpu():
    """``bench.py --gpus N`` with N > 1 exits with status 2 and says why, for
    every workload, before touching a device."""
    import json
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for wl in ("c4", "c3", "c5"):
        p = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2",
                            "--workload", wl], capture_output=True, text=True, timeout=120)
        assert p.returncode == 2, (wl, p.returncode, p.stderr[-500:])
        msg = json.loads(p.stdout.strip().splitlines()[-1])
        assert "one GPU" in msg["error"]


class _FakeLoop:
    """Stand-in for the device time loop with the same contract: x_k =
    0.5 x_{k-1} + sum_s scale[k][s] src[s] on free nodes, Dirichlet nodes at
    their group's value; stored every store_every steps from x_0; probe
    'current' = sum of x_k - x_{k-1} over all nodes (row 0 copies row 1)."""

    def __init__(self, group):
        self.group = group
        self.calls = 0

    def transient(self, Aexp, x0, nsteps, group=None, values=None, src=None, src_scale=None,
                  store_every=0, x_before=None, nprobe=0, **kw):
        self.calls += 1
        x = np.array(x0, dtype=float)
        xs, prev = [x.copy()], x.copy()
        pi = np.zeros((nsteps + 1, 1))
        ss = np.zeros((nsteps + 1, 4))
        ss[0] = [x.sum(), x.max(), 0, 0]
        for k in range(1, nsteps + 1):
            x = 0.5 * prev
            if src is not None:
                x = x + src_scale[k] @ np.asarray(src)
            if values is not None:
                fixed = group >= 0
                x[fixed] = values[k][group[fixed]]
            pi[k, 0] = float((x - prev).sum())
            ss[k] = [x.sum(), x.max(), 0, 0]
            xs.append(x.copy())
            prev_before, prev = prev, x
        pi[0] = pi[1]
        stored = np.array(xs[::store_every]) if store_every > 0 else None
        st = {k: 0 for k in ("iterations", "refreshes", "fallbacks", "pc_applies",
                             "replacements", "stop_reason")}
        st.update(residual=0.0, residual_eq=0.0, solve_time=0.0, converged=1)
        return {"x": x, "x_before": prev_before, "stored": stored, "probe_i": pi,
                "probe_v": None, "step_stats": ss, "stats": st, "rc": 0}


@pytest.mark.parametrize("store_every", [1, 3, 4])
def test_chunked_time_loop_equals_one_call(monkeypatch, store_every):
    """Sources that take more shapes than one device call holds, and a values
    array past the budget, split the run into calls; the stitched fields,
    probes and stats equal one call's."""
    from fieldspice_amd.solvers import transient as T
    rng = np.random.default_rng(7)
    n, nsteps = 12, 23
    group = np.full(n, -1, np.int32)
    group[[0, 5]] = [0, 1]
    values = rng.standard_normal((nsteps + 1, 2))
    sources = [None] + [rng.standard_normal(n) if k % 4 else None for k in range(1, nsteps + 1)]
    x0 = rng.standard_normal(n)
    one = _FakeLoop(group)
    monkeypatch.setattr(T, "MAX_SOURCE_BASIS", 64)
    ref = T.device_time_loop(one, None, x0, nsteps, group=group, values=values, sources=sources,
                             store_every=store_every)
    assert one.calls == 1
    many = _FakeLoop(group)
    monkeypatch.setattr(T, "MAX_SOURCE_BASIS", 3)
    monkeypatch.setattr(T, "VALUES_BUDGET_BYTES", 8 * 2 * 6)
    out = T.device_time_loop(many, None, x0, nsteps, group=group, values=values,
                             sources=sources, store_every=store_every)
    assert many.calls > 3
    for key in ("x", "stored", "probe_i", "step_stats"):
        np.testing.assert_allclose(out[key], ref[key], rtol=1e-13, atol=1e-13)
    assert out["stored"].shape[0] == nsteps // store_every + 1


def test_array_valued_wall_groups():
    """A wall with a value per node becomes one group per node; later claims
    still win (walls after terminals)."""
    g = RectilinearGrid.uniform([(0, 1.0), (0, 1.0), (0, 1.0)], [4, 3, 2])
    nid = g.node_ids()
    wall = wall_nodes(g, "xlo")
    prof = np.linspace(0.0, 1.0, wall.size)
    bc = BoundarySpec(xlo=Dirichlet(lambda t: (1.0 + t) * prof), xhi=Dirichlet(0.5))
    term = Terminal("t", nid[0:2, 0, 0].ravel(), voltage=2.0)
    G = DirichletGroups(g, [term], bc, walls_last=True)
    for t in (0.0, 0.3):
        idx, vals = G.node_values(t)
        w_idx, w_val = bc.dirichlet_nodes(g, t)
        ri, rv = O.unique_last(np.concatenate([term.nodes, w_idx]),
                               np.concatenate([np.full(term.nodes.size, 2.0), w_val]))
        assert np.array_equal(idx, ri) and np.array_equal(vals, rv)


def test_check_meshes_are_graded_and_closed():
    """The closed-form checks' meshes: fine and uniform inside the feature,
    growth at most the requested ratio (except a shorter closing cell), cells
    capped, endpoints exact; the junction mesh is finest at x = 0."""
    from fieldspice_amd.solvers import checks
    for pad in (2.0, 3.0, 5.0, 9.0):
        R, a = pad * 1e-6, 1e-6
        ax = checks._graded_axis(R, a, a / 14, max(a, R / 6), 1.25)
        d = np.diff(ax)
        assert ax[0] == 0.0 and ax[-1] == R and np.all(d > 0)
        assert np.allclose(d[:14], a / 14) and ax[14] == pytest.approx(a)
        assert d.max() <= max(a, R / 6) * (1 + 1e-12)
        assert np.all(d[1:-2] / d[:-3] <= 1.25 + 1e-12)
    x = checks._junction_axis(3e-7, 5e-7, 1e-9, 4e-8, 1.25)
    assert x[0] == pytest.approx(-3e-7) and x[-1] == pytest.approx(5e-7)
    assert 0.0 in x and np.all(np.diff(x) > 0)
    k = int(np.flatnonzero(x == 0.0)[0])
    assert np.diff(x)[k - 1] == pytest.approx(1e-9) and np.diff(x)[k] == pytest.approx(1e-9)
