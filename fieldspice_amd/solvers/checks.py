"""Closed-form self-checks of the Poisson path (the checks of
``fieldspice/solvers/poisson.py:1633-1987``), run on the device.

Each builds a small problem whose answer is known in closed form, solves it
with the GPU solvers and returns the measured numbers; the pass/fail bars
live in the tests.  They double as quick device smoke cases:

* a parallel plate spanning the cross-section (Neumann side walls: no
  fringing, so the discrete C is eps A / d to rounding);
* two dielectrics in series with the interface on a node plane (series
  capacitor formula and the divider potential at the interface);
* an isolated sphere (octant symmetry, grounded outer walls) against
  4 pi eps a as the box grows: the open-boundary truncation error, which
  converges from above like 4 pi eps a / (1 - a/R);
* an abrupt pn junction whose built-in potential is
  Vt [asinh(Na/2ni) + asinh(Nd/2ni)] (Vt ln(Na Nd / ni^2) to 1e-15 at these
  dopings), with one contact held and the other end free.
"""

from __future__ import annotations

from typing import Any, Sequence

import numpy as np

from ..boundaries import BoundarySpec, Dirichlet, Neumann
from ..grid import RectilinearGrid
from ..units import eps0, kB, q, thermal_voltage
from .base import SolverConfig, Terminal
from .poisson import NonlinearPoissonSolver, PoissonSolver, boltzmann_equilibrium_charge

__all__ = ["verify_parallel_plate", "verify_series_stack", "sphere_capacitance_convergence",
           "verify_built_in_potential", "verify_built_in_potential_sweep", "self_test",
           "NI_SI_300"]

NI_SI_300 = 9.65e15
"""Silicon intrinsic carrier density at 300 K [m^-3] (Sproul and Green, 1991)."""


def _plates(grid, axis=0):
    """Terminals on the two faces normal to ``axis``: low at 0 V, high at 1 V."""
    nid = grid.node_ids()
    return (Terminal("lo", np.take(nid, 0, axis=axis).ravel(), voltage=0.0),
            Terminal("hi", np.take(nid, -1, axis=axis).ravel(), voltage=1.0))


def verify_parallel_plate(gap: float = 2e-6, wy: float = 5e-6, wz: float = 4e-6,
                          eps_r: float = 3.9, n=(24, 10, 8),
                          config: SolverConfig | None = None) -> dict[str, Any]:
    """Plates across the whole cross-section: ``C = eps_r eps0 wy wz / gap``;
    the potential is uniform on every plane of constant x."""
    grid = RectilinearGrid.uniform([(0, gap), (0, wy), (0, wz)], n)
    ps = PoissonSolver(grid, np.full(grid.shape_cells, eps_r * eps0), config)
    lo, hi = _plates(grid)
    rep: dict[str, Any] = {}
    C = ps.capacitance_matrix([lo, hi], report=rep)
    C_ref = eps_r * eps0 * wy * wz / gap
    phi = ps.solve([lo, hi]).fields["phi"][0]
    return {"C": float(C[1, 1]), "C_ref": C_ref, "rel_err": abs(float(C[1, 1]) / C_ref - 1.0),
            "report": rep, "uniform_dev": float(np.abs(phi - phi[:, :1, :1]).max()),
            "n_nodes": grid.n_nodes}


def verify_series_stack(d1: float = 1.2e-6, d2: float = 0.8e-6, eps_r1: float = 3.9,
                        eps_r2: float = 7.5, wy: float = 5e-6, wz: float = 4e-6, n1: int = 12,
                        n2: int = 8, config: SolverConfig | None = None) -> dict[str, Any]:
    """Two layers in series, interface on a node plane:
    ``1/C = d1/(eps1 A) + d2/(eps2 A)``; the interface potential is the
    divider ``(d1/eps1) / (d1/eps1 + d2/eps2)`` (the low plate at 0 V)."""
    x = np.concatenate([np.linspace(0.0, d1, n1 + 1), np.linspace(d1, d1 + d2, n2 + 1)[1:]])
    grid = RectilinearGrid(x, np.linspace(0, wy, 5), np.linspace(0, wz, 4))
    eps = np.empty(grid.shape_cells)
    eps[:n1], eps[n1:] = eps_r1 * eps0, eps_r2 * eps0
    ps = PoissonSolver(grid, eps, config)
    lo, hi = _plates(grid)
    C = ps.capacitance_matrix([lo, hi])
    area = wy * wz
    C_ref = 1.0 / (d1 / (eps_r1 * eps0 * area) + d2 / (eps_r2 * eps0 * area))
    phi = ps.solve([lo, hi]).fields["phi"][0]
    v_if = float(phi[n1, 0, 0])
    v_ref = (d1 / eps_r1) / (d1 / eps_r1 + d2 / eps_r2)
    return {"C": float(C[1, 1]), "C_ref": C_ref, "rel_err": abs(float(C[1, 1]) / C_ref - 1.0),
            "v_interface": v_if, "v_interface_ref": v_ref, "v_rel_err": abs(v_if / v_ref - 1.0)}


def _graded_axis(extent, fine_until, dx_min, dx_max, growth):
    """Nodes of [0, extent]: uniform at ``dx_min`` up to ``fine_until``, then
    cells growing by ``growth`` per cell, capped at ``dx_max``; a remainder
    under 1.5 cells closes the axis in one or two cells of at most
    ``dx_max``."""
    n_fine = max(1, int(round(fine_until / dx_min)))
    nodes = list(np.linspace(0.0, fine_until, n_fine + 1))
    h = dx_min
    while True:
        h = min(h * growth, dx_max)
        rest = extent - nodes[-1]
        if rest <= 1.5 * h:
            if rest > dx_max:
                nodes.append(nodes[-1] + 0.5 * rest)
            nodes.append(extent)
            return np.array(nodes)
        nodes.append(nodes[-1] + h)


def sphere_capacitance_convergence(radius: float = 1e-6, pads: Sequence[float] = (2.0, 3.0, 5.0, 9.0),
                                   cells_per_radius: int = 14, eps_r: float = 1.0,
                                   growth: float = 1.25,
                                   config: SolverConfig | None = None) -> dict[str, Any]:
    """Isolated sphere in an octant box ``[0, pad a]^3`` (symmetry walls at 0,
    grounded walls at pad a), eight times the octant capacitance against
    ``4 pi eps a``: the truncation error of a bounded domain, which the
    concentric-shield value ``4 pi eps a / (1 - a/R)`` predicts to leading
    order (``deshielded = C (1 - a/R)`` removes it)."""
    C_ref = 4.0 * np.pi * eps_r * eps0 * radius
    out: dict[str, Any] = {"pads": tuple(float(p) for p in pads), "C_ref": C_ref}
    Cs, errs, shells, desh, nn = [], [], [], [], []
    for pad in pads:
        R = float(pad) * radius
        ax = _graded_axis(R, radius, radius / cells_per_radius, max(radius, R / 6.0), growth)
        grid = RectilinearGrid(ax, ax, ax)
        ps = PoissonSolver(grid, np.full(grid.shape_cells, eps_r * eps0), config)
        X, Y, Z = grid.node_coords()
        inside = (X ** 2 + Y ** 2 + Z ** 2) <= radius ** 2 * (1.0 + 1e-12)
        term = Terminal("sphere", np.flatnonzero(inside.ravel()), voltage=1.0)
        bc = BoundarySpec(xhi=Dirichlet(0.0), yhi=Dirichlet(0.0), zhi=Dirichlet(0.0),
                          xlo=Neumann(), ylo=Neumann(), zlo=Neumann())
        C_full = 8.0 * float(ps.capacitance_matrix([term], bc)[0, 0])
        Cs.append(C_full)
        errs.append(C_full / C_ref - 1.0)
        shells.append(C_ref / (1.0 - 1.0 / pad))
        desh.append(C_full * (1.0 - 1.0 / pad))
        nn.append(grid.n_nodes)
    out.update(C=tuple(Cs), rel_err=tuple(errs), shell_pred=tuple(shells),
               deshielded=tuple(desh), n_nodes=tuple(nn))
    return out


def _junction_axis(Lp, Ln, dx_min, dx_max, growth):
    """Nodes from -Lp to Ln, finest at the junction x = 0, growing away."""
    right = _graded_axis(Ln, dx_min, dx_min, dx_max, growth)
    left = _graded_axis(Lp, dx_min, dx_min, dx_max, growth)
    return np.concatenate([-left[::-1], right[1:]])


def verify_built_in_potential(Na: float = 1e23, Nd: float = 1e23, ni: float | None = None,
                              T: float = 300.0, eps_r: float = 11.7,
                              debye_multiple: float = 45.0, cells_per_debye: float = 10.0,
                              config: SolverConfig | None = None) -> dict[str, Any]:
    """Abrupt silicon pn junction: the p contact held at its charge-neutral
    bulk potential ``-Vt asinh(Na/2ni)``, every other wall Neumann; the
    built-in potential is ``phi(+Ln) - phi(-Lp)``, against ``Vt ln(Na Nd /
    ni^2)`` and the exact neutral-bulk difference ``Vt [asinh(Na/2ni) +
    asinh(Nd/2ni)]``."""
    ni = float(NI_SI_300 if ni is None else ni)
    Vt = thermal_voltage(T)
    debye = lambda N: float(np.sqrt(eps_r * eps0 * kB * T / (q * q * N)))  # noqa: E731
    Lp, Ln = debye_multiple * debye(Na), debye_multiple * debye(Nd)
    x = _junction_axis(Lp, Ln, min(debye(Na), debye(Nd)) / cells_per_debye,
                       max(Lp, Ln) / 15.0, 1.25)
    grid = RectilinearGrid(x)
    X, _, _ = grid.node_coords()
    eq = boltzmann_equilibrium_charge(np.where(X > 0.0, Nd, -Na).ravel(), ni, T)
    cfg = config or SolverConfig(newton_tol=1e-10)
    ns = NonlinearPoissonSolver(grid, np.full(grid.shape_cells, eps_r * eps0), eq.rho,
                                eq.drho_dphi, cfg, v_scale=eq.Vt)
    phi_p = -Vt * float(np.arcsinh(Na / (2.0 * ni)))
    res = ns.solve(bc=BoundarySpec(xlo=Dirichlet(phi_p)), phi0=eq.phi_guess)
    phi = res.fields["phi"][0]
    Vbi = float(phi[-1, 0, 0] - phi[0, 0, 0])
    Vbi_ref = float(Vt * np.log(Na * Nd / (ni * ni)))
    Vbi_exact = Vt * float(np.arcsinh(Na / (2.0 * ni)) + np.arcsinh(Nd / (2.0 * ni)))
    newton = res.meta["newton"]
    return {"Vbi": Vbi, "Vbi_ref": Vbi_ref, "rel_err": abs(Vbi / Vbi_ref - 1.0),
            "Vbi_exact": Vbi_exact, "rel_err_exact": abs(Vbi / Vbi_exact - 1.0),
            "iterations": newton["iterations"], "residual_history": newton["residual_history"],
            "damping": newton["damping"],
            "net_charge": float(np.sum(eq.rho(phi.ravel()) * grid.node_volumes().ravel())),
            "n_nodes": grid.n_nodes, "uniform_dev": float(np.abs(phi - phi[:, :1, :1]).max())}


def verify_built_in_potential_sweep(config: SolverConfig | None = None) -> dict[str, Any]:
    """The doping / ni sweep of the reference's contract: ni in {9.65e15,
    1e16, 1.45e16} m^-3 against (Na, Nd) in {(1e17, 1e17), (1e15, 1e18),
    (1e19, 1e16), (5e16, 3e17)} cm^-3."""
    cases = []
    for ni in (9.65e15, 1.0e16, 1.45e16):
        for na, nd in ((1e17, 1e17), (1e15, 1e18), (1e19, 1e16), (5e16, 3e17)):
            r = verify_built_in_potential(na * 1e6, nd * 1e6, ni=ni, config=config)
            cases.append({"ni": ni, "Na_cm3": na, "Nd_cm3": nd, "Vbi": r["Vbi"],
                          "Vbi_ref": r["Vbi_ref"], "rel_err": r["rel_err"],
                          "rel_err_exact": r["rel_err_exact"], "iterations": r["iterations"],
                          "n_nodes": r["n_nodes"]})
    return {"cases": cases, "max_rel_err": max(c["rel_err"] for c in cases),
            "max_rel_err_exact": max(c["rel_err_exact"] for c in cases),
            "max_iterations": max(c["iterations"] for c in cases)}


def self_test(verbose: bool = True) -> dict[str, Any]:
    """Every check above; the numbers are returned (and printed), not asserted."""
    out = {"parallel_plate": verify_parallel_plate(), "series_stack": verify_series_stack(),
           "pn_junction": verify_built_in_potential(),
           "sphere": sphere_capacitance_convergence()}
    if verbose:
        pp, ss, pn, sph = (out[k] for k in ("parallel_plate", "series_stack", "pn_junction",
                                             "sphere"))
        print(f"parallel plate : C = {pp['C']:.6e} F, rel err {pp['rel_err']:.2e}")
        print(f"series stack   : C = {ss['C']:.6e} F, rel err {ss['rel_err']:.2e}")
        print(f"pn junction    : Vbi = {pn['Vbi']:.9f} V, rel err {pn['rel_err']:.2e}, "
              f"{pn['iterations']} Newton iterations")
        for pad, C, e in zip(sph["pads"], sph["C"], sph["rel_err"]):
            print(f"sphere pad {pad:4.1f} a : C = {C:.6e} F, rel err {e:+.4f}")
    return out
