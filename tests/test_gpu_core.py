"""GPU parity of the core path against the CPU oracle (needs an MI355X).

Each test prints the case it runs before touching the device.
"""

import warnings

import numpy as np
import pytest

import fs_oracle as O
from fieldspice_amd import _lib, workloads
from fieldspice_amd.grid import RectilinearGrid, graded_1d
from fieldspice_amd.solvers.base import SolverConfig, Terminal
from fieldspice_amd.solvers.linear import GridLinearSystem
from fieldspice_amd.solvers.poisson import PoissonSolver
from fieldspice_amd.units import eps0

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[gpu case] {msg}", flush=True)


def _graded_grid():
    return RectilinearGrid(graded_1d(0, 3e-6, 13, 1.15), graded_1d(0, 2e-6, 9, 0.9),
                           graded_1d(0, 2.5e-6, 11, 1.2))


def test_device_is_gfx950(gpu_lib):
    _case("device info")
    name, ncu = _lib.device_info()
    assert name.startswith("gfx950"), name
    assert ncu >= 200


@pytest.mark.parametrize("mode", ["parallel", "harmonic", "min", "max"])
def test_edge_weights_match_oracle(gpu_lib, mode):
    _case(f"cell_to_edge+edge_mass mode={mode}")
    g = _graded_grid()
    rng = np.random.default_rng(1)
    prop = rng.uniform(1.0, 9.0, g.shape_cells) * eps0
    op = _lib.DeviceOperator.from_cells(g, prop, mode)
    ref = O.edge_weights(g, O.cell_to_edge(g, prop, mode))
    got = op.edges()
    assert np.abs(got - ref).max() <= 1e-15 * np.abs(ref).max()


def test_apply_matches_csr(gpu_lib):
    _case("raw stencil apply vs CSR")
    g = _graded_grid()
    rng = np.random.default_rng(2)
    prop = rng.uniform(1.0, 9.0, g.shape_cells)
    extra = rng.uniform(0.0, 1.0, g.n_nodes)
    w = O.edge_weights(g, O.cell_to_edge(g, prop))
    L = O.nodal_laplacian(g, O.cell_to_edge(g, prop))
    x = rng.standard_normal(g.n_nodes)
    op = _lib.DeviceOperator.from_edges(g, w, extra)
    y = op.apply(x)
    ref = L @ x + extra * x
    assert np.abs(y - ref).max() <= 1e-13 * np.abs(ref).max()
    # combine: 2*A + 0.5*B + diag(3*e)
    opB = _lib.DeviceOperator.from_cells(g, prop)
    e2 = rng.uniform(0.0, 1.0, g.n_nodes)
    comb = op.combine(2.0, 0.5, opB, 3.0, e2)
    ref2 = 2.0 * ref + 0.5 * (L @ x) + 3.0 * e2 * x
    assert np.abs(comb.apply(x) - ref2).max() <= 1e-13 * np.abs(ref2).max()


def _dirichlet_problem(seed=3):
    g = _graded_grid()
    rng = np.random.default_rng(seed)
    prop = rng.choice([2.7, 3.9, 7.5], size=g.shape_cells) * eps0
    nid = g.node_ids()
    fixed = np.concatenate([nid[0].ravel(), nid[-1].ravel(), nid[5:8, 3:6, 4:7].ravel()])
    vals = np.concatenate([np.full(nid[0].size, 1.0), np.zeros(nid[-1].size),
                           np.full(27, 0.3)])
    fixed, vals = O.unique_last(fixed, vals)
    L = O.nodal_laplacian(g, O.cell_to_edge(g, prop))
    rhs = rng.standard_normal(g.n_nodes) * 1e-18
    A, b = O.apply_dirichlet(L, rhs, fixed, vals)
    return g, prop, fixed, vals, A, b


@pytest.mark.parametrize("solver", ["hip_pcg", "hip_mixed"])
def test_general_solve_matches_direct(gpu_lib, solver):
    _case(f"general Dirichlet solve {solver}")
    g, prop, fixed, vals, A, b = _dirichlet_problem()
    x_ref = O.LinearSystem(A, method="direct").solve(b)
    op = _lib.DeviceOperator.from_cells(g, prop)
    ls = GridLinearSystem(op, fixed, SolverConfig(linear_solver=solver))
    x = ls.solve(b)
    assert np.abs(x - x_ref).max() <= 1e-9 * np.abs(x_ref).max()
    assert np.all(x[fixed] == vals)
    assert ls.info["residual_eq"] < 1e-9


def test_iteration_count_matches_scipy_cg(gpu_lib):
    _case("iteration count vs scipy cg (Jacobi branch)")
    g, prop, fixed, vals, A, b = _dirichlet_problem()
    ols = O.LinearSystem(A, method="cg_jacobi", tol=1e-10)
    ols.solve(b)
    op = _lib.DeviceOperator.from_cells(g, prop)
    ls = GridLinearSystem(op, fixed, SolverConfig(linear_solver="hip_pcg"))
    ls.solve(b)
    assert abs(ls.info["iterations"] - ols.info["iterations"]) <= 2


@pytest.mark.parametrize("solver", ["hip_pcg", "hip_mixed"])
def test_initial_guess_solve(gpu_lib, solver):
    _case(f"general solve from an initial guess, {solver}")
    g, prop, fixed, vals, A, b = _dirichlet_problem()
    x_ref = O.LinearSystem(A, method="direct").solve(b)
    op = _lib.DeviceOperator.from_cells(g, prop)
    ls = GridLinearSystem(op, fixed, SolverConfig(linear_solver=solver))
    ls.solve(b)
    cold = ls.info["iterations"]
    # the exact solution: converged before the first iteration
    x = ls.solve(b, x0=x_ref)
    assert ls.info["iterations"] == 0
    assert np.abs(x - x_ref).max() <= 1e-9 * np.abs(x_ref).max()
    # a guess 1e-4 off on the free rows (Dirichlet entries of x0 are ignored:
    # the values come from b): fewer iterations, the same answer
    rng = np.random.default_rng(4)
    x0 = x_ref * (1.0 + 1e-4 * rng.standard_normal(x_ref.size))
    x0[fixed] = -7.0
    x = ls.solve(b, x0=x0)
    warm = ls.info["iterations"]
    assert warm < cold
    assert np.abs(x - x_ref).max() <= 1e-9 * np.abs(x_ref).max()
    assert np.all(x[fixed] == vals)
    if solver == "hip_pcg":
        # the reference's CG from the same guess (scipy cg, x0 scaled as
        # fieldspice/solvers/poisson.py:543-547) takes as many iterations
        ols = O.LinearSystem(A, method="cg_jacobi", tol=1e-10)
        ols.solve(b, x0=x0)
        assert abs(warm - ols.info["iterations"]) <= 2
    # two right-hand sides, one with a guess per row
    X = ls.solve(np.stack([b, b]), x0=np.stack([x0, np.zeros_like(x0)]))
    for k in range(2):
        assert np.abs(X[k] - x_ref).max() <= 1e-9 * np.abs(x_ref).max()
    with pytest.raises(ValueError, match="x0"):
        ls.solve(b, x0=x0[:-1])
    with pytest.raises(ValueError, match="x0"):
        ls.solve(b, x0=np.full_like(x0, np.nan))


def test_multi_rhs_solve(gpu_lib):
    _case("multi-RHS general solve, 5 right-hand sides")
    g, prop, fixed, vals, A, b = _dirichlet_problem()
    rng = np.random.default_rng(9)
    B = np.stack([b * (1 + 0.1 * k) + rng.standard_normal(b.size) * 1e-19 * (k % 2)
                  for k in range(5)])
    for k in range(5):
        B[k, fixed] = vals
    op = _lib.DeviceOperator.from_cells(g, prop)
    ls = GridLinearSystem(op, fixed, SolverConfig())
    X = ls.solve(B)
    lu = O.LinearSystem(A, method="direct")
    for k in range(5):
        xr = lu.solve(B[k])
        assert np.abs(X[k] - xr).max() <= 1e-9 * np.abs(xr).max()


@pytest.mark.parametrize("solver", ["hip_pcg", "hip_mixed"])
def test_parallel_plate_capacitance(gpu_lib, solver):
    _case(f"c1 plate capacitance {solver}")
    ex = workloads.parallel_plate()
    ps = PoissonSolver(ex.grid, ex.eps_cell, SolverConfig(linear_solver=solver))
    rep = {}
    C = ps.capacitance_matrix(ex.terminals, report=rep)
    ref = 3.4531332486813394e-16
    assert C[0, 0] == pytest.approx(ref, rel=1e-10)
    assert C[1, 1] == pytest.approx(ref, rel=1e-10)
    assert C[0, 1] == pytest.approx(-ref, rel=1e-10)
    assert rep["row_sum_rel"] < 1e-9 and rep["asymmetry_rel"] < 1e-9


@pytest.mark.parametrize("solver", ["hip_pcg", "hip_mixed", "hip_fcg"])
def test_multiconductor_capacitance_matches_oracle(gpu_lib, solver):
    _case(f"layered 6-box capacitance {solver}")
    ex = workloads.layered_boxes(nodes=(34, 30, 22), n_boxes=6, side=(2, 6), seed=4)
    Cr, _, phis = O.capacitance_matrix(ex.grid, ex.eps_cell, [t.nodes for t in ex.terminals],
                                       method="direct", return_phi=True)
    ps = PoissonSolver(ex.grid, ex.eps_cell, SolverConfig(linear_solver=solver))
    C, phi = ps.capacitance_matrix(ex.terminals, return_phi=True)
    assert np.abs(C - Cr).max() <= 1e-9 * np.abs(Cr).max()
    for j in range(len(ex.terminals)):
        assert np.abs(phi[j] - phis[j]).max() <= 1e-9
    assert np.all(np.diag(C) > 0) and np.all(C - np.diag(np.diag(C)) <= 0)


def test_capacitance_with_grounded_shield(gpu_lib):
    _case("capacitance with Dirichlet shield walls")
    from fieldspice_amd.boundaries import BoundarySpec, Dirichlet
    ex = workloads.layered_boxes(nodes=(30, 26, 20), n_boxes=3, side=(2, 5), seed=7)
    bc = BoundarySpec(xlo=Dirichlet(0.0), zhi=Dirichlet(0.0))
    wall, _ = bc.dirichlet_nodes(ex.grid)
    Cr, rr = O.capacitance_matrix(ex.grid, ex.eps_cell, [t.nodes for t in ex.terminals],
                                  wall_fixed=wall, method="direct")
    rep = {}
    C = PoissonSolver(ex.grid, ex.eps_cell).capacitance_matrix(ex.terminals, bc, report=rep)
    assert np.abs(C - Cr).max() <= 1e-9 * np.abs(Cr).max()
    assert np.abs(rep["shield_charge"] - rr["shield_charge"]).max() <= 1e-9 * np.abs(Cr).max()


def test_more_than_eight_terminals(gpu_lib):
    _case("capacitance with 10 terminals (two batches)")
    ex = workloads.layered_boxes(nodes=(44, 40, 24), n_boxes=10, side=(2, 4), seed=11)
    Cr, _ = O.capacitance_matrix(ex.grid, ex.eps_cell, [t.nodes for t in ex.terminals],
                                 method="direct")
    C = PoissonSolver(ex.grid, ex.eps_cell).capacitance_matrix(ex.terminals)
    assert np.abs(C - Cr).max() <= 1e-9 * np.abs(Cr).max()


def test_floating_system_pins_gauge_like_reference(gpu_lib):
    _case("singular all-Neumann system gauge pin")
    g = _graded_grid()
    rng = np.random.default_rng(5)
    prop = rng.uniform(1, 3, g.shape_cells)
    L = O.nodal_laplacian(g, O.cell_to_edge(g, prop))
    with pytest.warns(RuntimeWarning):
        ols = O.LinearSystem(L, method="direct")
    op = _lib.DeviceOperator.from_cells(g, prop)
    with pytest.warns(RuntimeWarning, match="gauge"):
        ls = GridLinearSystem(op, None, SolverConfig())
    assert ls.info["pinned"] == ols.gauge
    b = rng.standard_normal(g.n_nodes)
    b -= b.mean()
    assert np.abs(ls.solve(b) - ols.solve(b)).max() <= 1e-8 * np.abs(ols.solve(b)).max()
    with pytest.raises(ValueError, match="singular"):
        GridLinearSystem(op, None, SolverConfig(), pin=False)


def test_empty_rows_are_pinned(gpu_lib):
    _case("empty rows (zero-conductivity island)")
    g = RectilinearGrid.uniform([(0, 1.0), (0, 1.0), (0, 1.0)], [6, 5, 4])
    sig = np.ones(g.shape_cells)
    sig[2:5, 1:4, 1:3] = 0.0   # interior node (3..4, 2..3, 2) sees only zero cells
    L = O.nodal_laplacian(g, O.cell_to_edge(g, sig))
    nid = g.node_ids()
    fixed = np.concatenate([nid[0].ravel(), nid[-1].ravel()])
    vals = np.concatenate([np.ones(nid[0].size), np.zeros(nid[0].size)])
    A, b = O.apply_dirichlet(L, np.zeros(g.n_nodes), fixed, vals)
    with pytest.warns(RuntimeWarning):
        x_ref = O.LinearSystem(A).solve(b)
    op = _lib.DeviceOperator.from_cells(g, sig)
    with pytest.warns(RuntimeWarning, match="empty rows"):
        ls = GridLinearSystem(op, fixed, SolverConfig())
    x = ls.solve(b)
    assert np.abs(x - x_ref).max() <= 1e-9


def test_forced_nonconvergence_raises(gpu_lib):
    _case("forced non-convergence (maxiter=2)")
    from fieldspice_amd.solvers.base import ConvergenceError
    g, prop, fixed, vals, A, b = _dirichlet_problem()
    op = _lib.DeviceOperator.from_cells(g, prop)
    ls = GridLinearSystem(op, fixed, SolverConfig(maxiter=2))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        with pytest.raises(ConvergenceError):
            ls.solve(b)


def test_nan_rhs_raises(gpu_lib):
    _case("NaN right-hand side")
    from fieldspice_amd.solvers.base import ConvergenceError
    g, prop, fixed, vals, A, b = _dirichlet_problem()
    b = b.copy()
    b[np.setdiff1d(np.arange(b.size), fixed)[7]] = np.nan
    op = _lib.DeviceOperator.from_cells(g, prop)
    ls = GridLinearSystem(op, fixed, SolverConfig())
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        with pytest.raises((ConvergenceError, ValueError)):
            ls.solve(b)
