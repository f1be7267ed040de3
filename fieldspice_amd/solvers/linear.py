"""GPU counterpart of fieldspice's ``LinearSystem`` for grid operators.

``fieldspice/solvers/poisson.py:159-589`` prepares a CSR matrix once
(empty-row and gauge pinning, Dirichlet-row split, Jacobi equilibration) and
solves it for many right-hand sides with SciPy CG, stopping at
``||r_s|| < tol*||b_s||`` in the equilibrated space.  ``GridLinearSystem`` does
the same preparation on the GPU from the operator's edge weights and a
Dirichlet mask, and runs the same CG iteration there.  Diagnostics keep the
reference's ``info`` keys; the reported ``residual`` is the reference's
``||D^1/2 (b_s - A_s x_s)|| / ||b||`` and ``residual_eq`` the equilibrated
``||b_s - A_s x_s|| / ||b_s||`` (both recomputed in FP64 on the device).
"""

from __future__ import annotations

import time
import warnings

import numpy as np

from .. import _lib
from .base import ConvergenceError, SolverConfig

__all__ = ["GridLinearSystem", "method_for"]


def _scipy_flag(st: dict) -> str:
    """The flag ``fieldspice/solvers/poisson.py:554-557`` prints: SciPy's
    ``cg`` returns its iteration cap as ``info`` whenever it stops without
    converging.  SciPy has no breakdown exit; the device stops on one
    (non-finite or non-positive curvature) and says so."""
    flag = str(int(st["maxiter"]))
    return flag + " (breakdown)" if st["stop_reason"] == 3 else flag


def method_for(cfg: SolverConfig) -> tuple[str, int]:
    """(method name, precision code) the GPU runs for ``cfg.linear_solver``."""
    want = cfg.linear_solver
    if want == "hip_mixed":
        return "hip_mixed", _lib.FS_PREC_MIXED
    if want == "hip_fcg":
        return "hip_fcg", _lib.FS_PREC_FCG
    return "hip_pcg", _lib.FS_PREC_FP64


class GridLinearSystem:
    """Prepared GPU solve of ``op`` with Dirichlet rows ``fixed``.

    Parameters
    ----------
    op : DeviceOperator
        ``G^T diag(w) G + diag(extra)`` on the grid.
    fixed : int array or bool mask, optional
        Dirichlet nodes (identity rows after elimination).
    config : SolverConfig
    pin : bool
        Pin the gauge of a floating system (warning) instead of raising.
    grid : RectilinearGrid, optional
        ``hip_fcg``: the node spacings that shape the preconditioner's
        hierarchy (uniform spacing is assumed without it).
    """

    def __init__(self, op, fixed=None, config: SolverConfig | None = None, pin: bool = True,
                 grid=None):
        t0 = time.perf_counter()
        self.cfg = config or SolverConfig()
        self.method, self.precision = method_for(self.cfg)
        n = op.n_nodes
        mask = np.zeros(n, dtype=np.uint8)
        if fixed is not None:
            f = np.asarray(fixed)
            if f.dtype == bool:
                mask[f.ravel()] = 1
            else:
                mask[f.astype(np.intp)] = 1
        if getattr(op, "periodic", 0) and (self.precision == _lib.FS_PREC_FCG
                                           or self.cfg.deflate_islands):
            raise ValueError("periodic walls run with hip_pcg or hip_mixed (no learned "
                             "preconditioner, no island space)")
        self.n = n
        self.op = op
        self.sys = _lib.DeviceSystem(op, mask, pin=pin)
        si = self.sys.info
        if si["n_empty"]:
            warnings.warn(f"{si['n_empty']} unknown(s) are coupled to nothing (empty rows); "
                          "they are pinned to zero.  In a conductance matrix this means a "
                          "region of zero conductivity that no terminal reaches.",
                          RuntimeWarning, stacklevel=2)
        if si["singular"]:
            warnings.warn("singular all-Neumann system: no Dirichlet node was found, so node "
                          f"{si['gauge']} is pinned to 0 to fix the gauge.  Only potential "
                          "DIFFERENCES in the result are meaningful; the absolute level is "
                          "arbitrary.", RuntimeWarning, stacklevel=2)
        self.singular = bool(si["singular"])
        self.n_free = int(si["n_free"])
        self.info = {
            "method": self.method,
            "requested": self.cfg.linear_solver,
            "n": n,
            "n_factorised": self.n_free,
            "n_decoupled": int(si["n_fixed"]),
            "equilibrated": True,
            "symmetric": True,
            "singular": self.singular,
            "pinned": int(si["gauge"]) if si["gauge"] >= 0 else None,
            "n_pinned": int(si["n_empty"]) + (1 if si["gauge"] >= 0 else 0),
            "setup_time": time.perf_counter() - t0,
            "notes": (),
        }
        self.pc = None
        self.info["islands"] = 0
        if self.cfg.deflate_islands:
            if self.precision == _lib.FS_PREC_FCG:
                raise ValueError("deflate_islands applies to hip_pcg and hip_mixed, not "
                                 "hip_fcg")
            ids, nisl = self.sys.find_islands(self.cfg.island_contrast)
            if nisl:
                self.sys.set_deflation(ids, nisl)
            self.info["islands"] = nisl
            self.info["setup_time"] = time.perf_counter() - t0
        if self.precision == _lib.FS_PREC_FCG:
            from ..precond import LearnedPreconditioner, uniform_widths
            lp = self.cfg.preconditioner or LearnedPreconditioner()
            widths = ((grid.ax.hd, grid.ay.hd, grid.az.hd) if grid is not None
                      else uniform_widths(op.shape_cells))
            self.pc = lp.attach_widths(self.sys, widths)
            self.info["precond_levels"] = self.pc.nlevels
            self.info["setup_time"] = time.perf_counter() - t0

    def params(self):
        return _lib.DeviceSystem.params(self.cfg.tol, self.cfg.maxiter, self.precision,
                                        self.cfg.inner_tol or 0.0, self.cfg.check_every)

    def _verdict(self, st: dict, rc: int, what: str = "cg"):
        """Convergence verdict of ``fieldspice/solvers/poisson.py:554-587``,
        applied to both residual norms."""
        self.info.update(iterations=st["iterations"], residual=st["residual"],
                         residual_eq=st["residual_eq"], refreshes=st["refreshes"],
                         solve_time=st["solve_time"], fallbacks=st["fallbacks"],
                         precond_applies=st["pc_applies"], replacements=st["replacements"],
                         stencil_width=st["stencil_width"], cycle_width=st["cycle_width"])
        if st["fallbacks"]:
            warnings.warn("flexible CG did not converge; the solve was redone with FP64 CG",
                          RuntimeWarning, stacklevel=3)
        if rc == _lib.FS_NOT_CONVERGED:
            warnings.warn(f"{self.method} returned flag {_scipy_flag(st)} after "
                          f"{st['iterations']} iterations", RuntimeWarning, stacklevel=3)
        worst = max(st["residual"], st["residual_eq"])
        tol = self.cfg.tol
        if worst > max(1e-6, 100.0 * tol):
            raise ConvergenceError(
                f"{self.method} failed: relative residual {worst:.3e} after "
                f"{st['iterations']} iterations (requested {tol:.1e})", history=[worst])
        if worst > max(tol * 10.0, 1e-12):
            warnings.warn(f"{self.method} stopped at relative residual {worst:.3e}, looser "
                          f"than the requested {tol:.1e}", RuntimeWarning, stacklevel=3)

    def solve(self, b, x0=None) -> np.ndarray:
        """``LinearSystem.solve``: ``b`` carries the Dirichlet values on
        Dirichlet rows; several right-hand sides may be stacked as (k, n).
        ``x0`` (the shape of ``b``) is an initial guess: its free entries,
        equilibrated, start the iteration (``fieldspice/solvers/poisson.py:543-547``);
        the Dirichlet values still come from ``b``."""
        b = np.asarray(b, dtype=np.float64)
        if b.shape[-1] != self.n or b.ndim > 2:
            raise ValueError(f"right-hand side must have shape ({self.n},), got {b.shape}")
        if x0 is not None:
            x0 = np.asarray(x0, dtype=np.float64)
            if x0.shape != b.shape:
                raise ValueError(f"x0 must have shape {b.shape}, got {x0.shape}")
            if not np.all(np.isfinite(x0)):
                raise ValueError("x0 must be finite")
        x, st, rc = self.sys.solve(b, self.params(), x0=x0)
        self._verdict(st, rc)
        return x
