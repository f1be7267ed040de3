"""Write golden input/output vectors from the fieldspice reference itself.

Run in the development container only (it imports the reference package from
/root/reference, which never travels with this repository):

    python tests/golden/make_golden.py

Covers what no closed form pins: Maxwell capacitance and potentials of a
graded, layered multi-conductor geometry (with and without a grounded shield),
cell_to_edge on graded multi-material grids, an EQS transient through a
two-phase medium, a thermal transient with convection, cell_to_node on graded
multi-material grids, the gauge node LinearSystem pins on singular
(all-Neumann) systems with and without ties in the diagonal, and a thermal
transient through a medium of two heat capacities.  Outputs are small .npz
files beside this script; inputs are regenerated from the stored seeds and
arrays.  ``python tests/golden/make_golden.py [set ...]`` writes only the named
sets (``main``, ``nodes``).
"""

from __future__ import annotations

import os
import sys
import warnings

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def main():
    sys.path.insert(0, REF)
    from fieldspice import operators as ops
    from fieldspice.boundaries import BoundarySpec, Dirichlet
    from fieldspice.grid import RectilinearGrid, graded_1d
    from fieldspice.solvers.base import SolverConfig, Terminal
    from fieldspice.solvers.eqs import EQSSolver
    from fieldspice.solvers.poisson import PoissonSolver
    from fieldspice.solvers.thermal import ThermalSolver
    from fieldspice.units import eps0

    rng = np.random.default_rng(20261203)
    # -- graded layered multi-conductor capacitance ------------------------
    x = graded_1d(0, 2.0e-6, 22, 1.04)
    y = np.linspace(0, 1.6e-6, 17)
    z = graded_1d(0, 1.2e-6, 14, 1.12)
    g = RectilinearGrid(x, y, z)
    eps_r = np.empty(g.shape_cells)
    eps_r[:, :, :5] = 3.9
    eps_r[:, :, 5:9] = 7.5
    eps_r[:, :, 9:] = 2.7
    boxes = [(2, 6, 2, 6, 1, 4), (9, 14, 8, 12, 6, 9), (16, 20, 3, 7, 10, 13)]
    nid = np.arange(g.n_nodes).reshape(g.shape_nodes)
    terms = [Terminal(f"c{m}", nid[a:b + 1, c:d + 1, e:f + 1].ravel())
             for m, (a, b, c, d, e, f) in enumerate(boxes)]
    ps = PoissonSolver(g, eps_r * eps0)
    rep = {}
    C = ps.capacitance_matrix(terms, report=rep)
    bc = BoundarySpec(xlo=Dirichlet(0.0), zhi=Dirichlet(0.0))
    rep_s = {}
    Cs = ps.capacitance_matrix(terms, bc, report=rep_s)
    drv = [Terminal("c0", terms[0].nodes, voltage=1.0), Terminal("c1", terms[1].nodes, voltage=0.0),
           Terminal("c2", terms[2].nodes, voltage=0.25)]
    res = ps.solve(drv)
    e_edge = ops.cell_to_edge(g, eps_r * eps0)
    e_edge_h = ops.cell_to_edge(g, eps_r * eps0, mode="harmonic")
    np.savez(os.path.join(HERE, "capacitance_graded_layered.npz"),
             x=x, y=y, z=z, eps_r=eps_r, boxes=np.array(boxes), C=C, C_shield=Cs,
             shield_charge=rep_s["shield_charge"], asymmetry_rel=rep["asymmetry_rel"],
             row_sum=rep["row_sum"], phi=res.fields["phi"][0].ravel(),
             drive=np.array([1.0, 0.0, 0.25]), eps_edge=e_edge, eps_edge_harmonic=e_edge_h)

    # -- EQS transient through a two-phase medium -------------------------
    g2 = RectilinearGrid(np.linspace(0, 1.2e-7, 13), np.linspace(0, 1.0e-7, 11),
                         np.linspace(0, 0.8e-7, 9))
    sig = np.zeros(g2.shape_cells)
    sig[1:4, 1:4, 1:4] = 5.8e7
    sig[8:11, 6:9, 4:7] = 5.8e7
    sig[3:9, 2:6, 6:8] = 2.0e4        # resistive bridge
    eps = np.full(g2.shape_cells, 3.9 * eps0)
    eps[:, :, :3] = 7.5 * eps0
    nid2 = np.arange(g2.n_nodes).reshape(g2.shape_nodes)
    f0 = 1e9
    t_drv = Terminal("drv", nid2[1:5, 1:5, 1:5].ravel(),
                     voltage=lambda t: np.sin(2 * np.pi * f0 * t))
    t_gnd = Terminal("gnd", nid2[8:12, 6:10, 4:8].ravel(), voltage=0.0)
    dt, nsteps = 2e-12, 30
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        r = EQSSolver(g2, eps, sig, SolverConfig(verbose=0)).solve(
            [t_drv, t_gnd], t_end=nsteps * dt, dt=dt, store=True, store_every=10)
    np.savez(os.path.join(HERE, "eqs_two_phase.npz"),
             x=g2.xn, y=g2.yn, z=g2.zn, eps=eps, sigma=sig, f0=f0, dt=dt, nsteps=nsteps,
             drv_box=np.array([1, 5, 1, 5, 1, 5]), gnd_box=np.array([8, 12, 6, 10, 4, 8]),
             phi=r.fields["phi"], t_stored=r.scalars["t_stored"],
             i_drv=r.i("drv"), i_gnd=r.i("gnd"), v_drv=r.v("drv"),
             energy=r.scalars["energy_electric"], dissipation=r.scalars["dissipation"])

    # -- thermal transient with convection --------------------------------
    g3 = RectilinearGrid(np.linspace(0, 4e-5, 11), np.linspace(0, 3e-5, 7),
                         graded_1d(0, 2e-5, 6, 1.3))
    kap = np.full(g3.shape_cells, 1.4)
    kap[3:7, 2:5, :3] = 148.0
    rcp = np.full(g3.shape_cells, 2200.0 * 730.0)
    q = np.zeros(g3.n_nodes)
    nid3 = np.arange(g3.n_nodes).reshape(g3.shape_nodes)
    q[nid3[4:7, 2:5, 0].ravel()] = 1e-4
    conv = {"zhi": (1e4, 300.0)}
    bcT = BoundarySpec(xlo=Dirichlet(300.0))
    ts = ThermalSolver(g3, kap, rcp)
    rT = ts.transient(q, 2e-6, 1e-7, T0=300.0, bc=bcT, convection=conv, store=True)
    rS = ts.steady(q, bc=bcT, convection=conv)
    np.savez(os.path.join(HERE, "thermal_convective.npz"),
             x=g3.xn, y=g3.yn, z=g3.zn, kappa=kap, rho_cp=rcp, q=q, T=rT.fields["T"],
             T_max=rT.scalars["T_max"], T_mean=rT.scalars["T_mean"],
             T_steady=rS.fields["T"][0], t_end=2e-6, dt=1e-7)
    print("golden fixtures written to", HERE)


def nodes_gauge_thermal():
    """cell_to_node on graded multi-material grids, gauge pins of singular
    systems, and a thermal transient with heterogeneous rho*c_p."""
    sys.path.insert(0, REF)
    from fieldspice import operators as ops
    from fieldspice.boundaries import BoundarySpec, Dirichlet
    from fieldspice.grid import RectilinearGrid, graded_1d
    from fieldspice.solvers.base import SolverConfig
    from fieldspice.solvers.poisson import LinearSystem
    from fieldspice.solvers.thermal import ThermalSolver

    rng = np.random.default_rng(20261204)
    # -- cell_to_node: graded axes, blocky materials plus random cells ------
    x = graded_1d(0, 3.0e-6, 9, 1.3)
    y = graded_1d(0, 2.0e-6, 7, 0.8)
    z = graded_1d(0, 1.5e-6, 5, 1.5)
    g = RectilinearGrid(x, y, z)
    prop = np.full(g.shape_cells, 2.0)
    prop[3:7, 1:4, :2] = 11.0
    prop[:, 5:, 3:] = 0.5
    prop[rng.random(g.shape_cells) < 0.2] = rng.uniform(20.0, 40.0)
    c2n = ops.cell_to_node(g, prop)
    # -- gauge pins: a heterogeneous graded Laplacian and a uniform one -----
    sig = rng.uniform(1.0, 3.0, g.shape_cells)
    L1 = ops.nodal_laplacian(g, ops.cell_to_edge(g, sig))
    b1 = rng.standard_normal(g.n_nodes)
    b1 -= b1.mean()
    gu = RectilinearGrid(np.linspace(0, 1.0, 7), np.linspace(0, 1.0, 6), np.linspace(0, 1.0, 5))
    L2 = ops.nodal_laplacian(gu, ops.cell_to_edge(gu, np.ones(gu.shape_cells)))
    b2 = rng.standard_normal(gu.n_nodes)
    b2 -= b2.mean()
    cfg = SolverConfig(verbose=0, linear_solver="direct")
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        ls1 = LinearSystem(L1, cfg)
        x1 = ls1.solve(b1)
        ls2 = LinearSystem(L2, cfg)
        x2 = ls2.solve(b2)
    # -- thermal transient, two heat capacities and two conductivities -------
    gt = RectilinearGrid(np.linspace(0, 4e-5, 13), graded_1d(0, 3e-5, 8, 1.2),
                         np.linspace(0, 2e-5, 6))
    kap = np.full(gt.shape_cells, 1.4)
    kap[4:9, 2:6, :3] = 148.0
    rcp = np.full(gt.shape_cells, 2200.0 * 730.0)
    rcp[4:9, 2:6, :3] = 2329.0 * 700.0
    rcp[:, :, 4:] = 8960.0 * 385.0
    q = np.zeros(gt.n_nodes)
    nid = np.arange(gt.n_nodes).reshape(gt.shape_nodes)
    q[nid[5:8, 3:5, 0].ravel()] = 2e-4
    bcT = BoundarySpec(xlo=Dirichlet(300.0))
    ts = ThermalSolver(gt, kap, rcp)
    rT = ts.transient(q, 3e-6, 1e-7, T0=300.0, bc=bcT, store=True)
    np.savez(os.path.join(HERE, "nodes_gauge_thermal.npz"),
             x=x, y=y, z=z, prop=prop, cell_to_node=c2n,
             sigma=sig, b1=b1, x1=x1, gauge1=ls1.info["pinned"],
             b2=b2, x2=x2, gauge2=ls2.info["pinned"],
             tx=gt.xn, ty=gt.yn, tz=gt.zn, kappa=kap, rho_cp=rcp, q=q, T=rT.fields["T"],
             T_max=rT.scalars["T_max"], T_mean=rT.scalars["T_mean"], t_end=3e-6, dt=1e-7,
             C_th=ops.cell_to_node(gt, rcp).ravel() * gt.node_volumes().ravel())
    print("node/gauge/thermal fixtures written to", HERE)


if __name__ == "__main__":
    sets = sys.argv[1:] or ["main", "nodes"]
    if "main" in sets:
        main()
    if "nodes" in sets:
        nodes_gauge_thermal()
