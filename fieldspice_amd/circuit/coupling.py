"""A quasi-static field region reduced onto its terminals on the GPU
(``fieldspice/circuit/coupling.py:56-296``).

With ``K = L_eps/dt + L_sigma`` (``L_sigma`` at DC) partitioned into electrode
(t) and interior (i) nodes, the region is the ``nE x nE`` admittance
``Y = K_tt - K_ti K_ii^-1 K_it`` plus, in a transient step, the history
current ``f_eff = f_t - K_ti K_ii^-1 f_i`` with ``f = (L_eps/dt) phi_prev``.
``Y[a, b]`` is the current into electrode ``a`` with electrode ``b`` at 1 V and
the others at 0 V, i.e. the unit-excitation charge integral of the
capacitance path with ``K`` in place of ``L_eps``: all ``nE`` interior solves
run as one multi-right-hand-side device solve and the electrode sums are
integrated on the device.  The history term and the interior back-substitution
of an accepted step are single device solves of the same system.

The AC stamp needs ``K = L_sigma + j omega L_eps``, complex symmetric and not
Hermitian, which CG does not solve.  Written as the real 2x2 block system
``[[B, -C], [C, B]] [x; y] = [f; g]`` (``B = L_sigma``, ``C = omega L_eps`` on
the interior), it is solved by flexible GMRES preconditioned with PRESB
(Axelsson, Neytcheva and Ahmad, Numer. Algorithms 66 (2014) 811-841): the
preconditioner ``[[B, -C], [C, B + 2C]]`` is applied exactly by two solves with
``H = B + C``, a real SPD grid Laplacian that the device solves like any
other, and bounds the preconditioned spectrum to ``[1/2, 1]`` for every
``omega``, so the iteration count does not grow with the conductivity
contrast or the frequency.
"""

from __future__ import annotations

from typing import Any, Mapping, Sequence

import numpy as np

from .. import _lib
from ..solvers.base import SolverConfig, Terminal
from ..solvers.linear import GridLinearSystem

__all__ = ["FieldRegion", "stamp_g", "stamp_i", "presb_fgmres"]


def stamp_g(G: Any, i: int, j: int, value: float) -> None:
    """``G[i, j] += value``; ground (index < 0) dropped; ``G`` may expose
    ``add(i, j, v)`` (``fieldspice/circuit/devices.py:186-200``)."""
    if i < 0 or j < 0:
        return
    add = getattr(G, "add", None)
    if add is not None:
        add(i, j, value)
    else:
        G[i, j] += value


def stamp_i(I: Any, i: int, value: float) -> None:
    """``I[i] += value``; ground dropped (``fieldspice/circuit/devices.py:203-207``)."""
    if i < 0:
        return
    I[i] += value


def presb_fgmres(apply_B, apply_C, solve_H, f, g, tol: float = 1e-10, restart: int = 30,
                 max_restarts: int = 10):
    """Solve ``(B + jC)(x + jy) = f + jg`` for real symmetric positive
    semi-definite ``B``, ``C`` with ``H = B + C`` positive definite.

    ``f``, ``g``: real ``(k, n)`` right-hand sides, solved together: every
    callable takes and returns a ``(k, n)`` array (``solve_H`` one multi-RHS
    solve).  Flexible GMRES on the real block form, right-preconditioned with
    PRESB; the columns iterate in lockstep, each with its own Hessenberg
    matrix, until every column's residual is below ``tol`` times its
    right-hand side.  Returns ``(x, y, info)``; raises ``RuntimeError`` when
    ``max_restarts`` cycles of ``restart`` iterations do not reach ``tol``.
    """
    f = np.asarray(f, dtype=np.float64)
    g = np.asarray(g, dtype=np.float64)
    k = f.shape[0]

    def A(u):                               # u: (k, 2, n)
        bx, by = apply_B(u[:, 0]), apply_B(u[:, 1])
        cx, cy = apply_C(u[:, 0]), apply_C(u[:, 1])
        return np.stack([bx - cy, cx + by], axis=1)

    def P_inv(v):                           # PRESB: two solves with H
        h = solve_H(v[:, 0] + v[:, 1])
        y = solve_H(apply_B(h) - v[:, 0])
        return np.stack([h - y, y], axis=1)

    def norms(u):
        return np.sqrt(np.einsum("kij,kij->k", u, u))

    rhs = np.stack([f, g], axis=1)
    bnorm = norms(rhs)
    target = tol * bnorm
    sol = np.zeros_like(rhs)
    r = rhs.copy()
    total = 0
    for cycle in range(max_restarts + 1):
        beta = norms(r)
        if np.all(beta <= target):
            return sol[:, 0], sol[:, 1], {"iterations": total, "restarts": cycle,
                                          "residual": float(np.max(beta / np.maximum(bnorm,
                                                                                    1e-300)))}
        if cycle == max_restarts:
            break
        m = restart
        V = [r / np.where(beta > 0, beta, 1.0)[:, None, None]]
        Z = []
        Hm = np.zeros((k, m + 1, m))
        cs = np.zeros((k, m))
        sn = np.zeros((k, m))
        s = np.zeros((k, m + 1))
        s[:, 0] = beta
        j_end = 0
        for j in range(m):
            z = P_inv(V[j])
            Z.append(z)
            w = A(z)
            for i in range(j + 1):          # modified Gram-Schmidt, per column
                hij = np.einsum("kij,kij->k", w, V[i])
                w -= hij[:, None, None] * V[i]
                Hm[:, i, j] = hij
            hn = norms(w)
            Hm[:, j + 1, j] = hn
            V.append(w / np.where(hn > 0, hn, 1.0)[:, None, None])
            for i in range(j):              # earlier rotations on the new column
                a, b = Hm[:, i, j].copy(), Hm[:, i + 1, j].copy()
                Hm[:, i, j] = cs[:, i] * a + sn[:, i] * b
                Hm[:, i + 1, j] = -sn[:, i] * a + cs[:, i] * b
            a, b = Hm[:, j, j], Hm[:, j + 1, j]
            rho = np.hypot(a, b)
            ok = rho > 0
            cs[:, j] = np.where(ok, a / np.where(ok, rho, 1.0), 1.0)
            sn[:, j] = np.where(ok, b / np.where(ok, rho, 1.0), 0.0)
            Hm[:, j, j] = rho
            Hm[:, j + 1, j] = 0.0
            s[:, j + 1] = -sn[:, j] * s[:, j]
            s[:, j] = cs[:, j] * s[:, j]
            j_end = j + 1
            total += 1
            if np.all(np.abs(s[:, j + 1]) <= 0.5 * target):
                break
        for c in range(k):                  # back substitution per column
            if beta[c] <= target[c]:
                continue
            Rm = Hm[c, :j_end, :j_end]
            d = np.abs(np.diag(Rm)) > 0
            jj = int(np.argmin(d)) if not np.all(d) else j_end
            yv = np.linalg.solve(Rm[:jj, :jj], s[c, :jj]) if jj else np.zeros(0)
            for i in range(jj):
                sol[c] += yv[i] * Z[i][c]
        r = rhs - A(sol)
    raise RuntimeError(
        f"PRESB-preconditioned GMRES did not reach relative residual {tol:.1e} in "
        f"{total} iterations (reached {float(np.max(norms(r) / np.maximum(bnorm, 1e-300))):.2e})")


class FieldRegion:
    """A meshed electroquasistatic region stamped into a netlist
    (``fieldspice/circuit/coupling.py:56-133``): each terminal is one circuit
    node; the region contributes R and C, no inductance."""

    linear = True
    dynamic = True
    assumptions = ("A1a", "A2", "A3", "A10", "A12")

    def __init__(self, name: str, grid, eps_cell, sigma_cell, terminals: Sequence[Terminal],
                 circuit_nodes: Sequence[str] | None = None, config: SolverConfig | None = None):
        self.name = str(name)
        self.grid = grid
        self.cfg = config or SolverConfig()
        self.terminals = list(terminals)
        if len(self.terminals) < 1:
            raise ValueError("a field region needs at least one terminal")
        if circuit_nodes is None:
            circuit_nodes = [t.circuit_node or t.name for t in self.terminals]
        circuit_nodes = [str(c) for c in circuit_nodes]
        if len(circuit_nodes) != len(self.terminals):
            raise ValueError("circuit_nodes must match terminals in length")
        self.nodes = tuple(circuit_nodes)
        eps_cell = np.asarray(eps_cell, dtype=float)
        sigma_cell = np.asarray(sigma_cell, dtype=float)
        for arr, nm in ((eps_cell, "eps_cell"), (sigma_cell, "sigma_cell")):
            if arr.shape != grid.shape_cells:
                raise ValueError(f"{nm} must have shape {grid.shape_cells}")
        if np.any(eps_cell <= 0):
            raise ValueError("eps_cell must be positive and ABSOLUTE [F/m]")
        n = grid.n_nodes
        flat = np.concatenate([t.nodes for t in self.terminals])
        if np.unique(flat).size != flat.size:
            raise ValueError("terminals overlap: a node belongs to two electrodes")
        self._flat = flat
        self._tid = np.full(n, -1, dtype=np.int32)
        for k, t in enumerate(self.terminals):
            self._tid[t.nodes] = k
        self._sigma_zero = not np.any(sigma_cell != 0.0)
        self.L_eps = _lib.DeviceOperator.from_cells(grid, eps_cell)
        self.L_sigma = _lib.DeviceOperator.from_cells(grid, sigma_cell)
        self._cache: dict[float, dict[str, Any]] = {}
        self._ac_cache: dict[float, np.ndarray] = {}
        self.last_iterations = 0
        self.ac_info: dict[str, Any] = {}
        self.phi = np.zeros(n)
        self.asymmetry: float = 0.0

    def extra_unknowns(self) -> tuple[str, ...]:
        return ()

    def _reduce(self, dt: float | None) -> dict[str, Any]:
        """Reduced admittance and the device system for one step size (cached)."""
        key = float(dt) if dt is not None else 0.0
        if key in self._cache:
            return self._cache[key]
        nE = len(self.terminals)
        if dt is None and self._sigma_zero:
            # an insulating region has no DC conductance: an open circuit
            entry = {"K": None, "Y": np.zeros((nE, nE)), "ls": None, "dt": dt}
            self.asymmetry = 0.0
            self._cache[key] = entry
            return entry
        K = self.L_sigma if dt is None else self.L_eps.combine(1.0 / dt, 1.0, self.L_sigma)
        ls = GridLinearSystem(K, self._tid >= 0, self.cfg, grid=self.grid)
        Y, _, _, st, rc = ls.sys.capacitance(K, self._tid, nE, ls.params())
        ls._verdict(st, rc)
        self.asymmetry = float(np.abs(Y - Y.T).max() / max(np.abs(Y).max(), 1e-300))
        entry = {"K": K, "Y": 0.5 * (Y + Y.T), "ls": ls, "dt": dt}
        self._cache[key] = entry
        return entry

    def _interior_solve(self, entry, f, v, guess=None) -> np.ndarray:
        """``phi`` with the electrodes at ``v`` and ``K_ii phi_i = f_i - K_it v``;
        ``guess`` (a full potential) starts the iteration."""
        K, ls = entry["K"], entry["ls"]
        x0 = np.zeros(self.grid.n_nodes)
        vals = np.concatenate([np.full(t.nodes.size, v[k]) for k, t in enumerate(self.terminals)])
        x0[self._flat] = vals
        b = f - K.apply(x0)
        b[self._flat] = vals
        return ls.solve(b, x0=guess)

    def _history(self, entry, dt) -> np.ndarray:
        """``f_eff`` [A]: the history current source of the previous step."""
        nE = len(self.terminals)
        if dt is None:
            return np.zeros(nE)
        f = self.L_eps.apply(self.phi) / dt
        f_t = np.array([float(f[t.nodes].sum()) for t in self.terminals])
        # u: interior solve with every electrode at 0; K_ti u = sum_T (K u)
        u = self._interior_solve(entry, f, np.zeros(nE))
        ku = entry["K"].apply(u)
        return f_t - np.array([float(ku[t.nodes].sum()) for t in self.terminals])

    def stamp_dc(self, G, I, x, nmap: Mapping[str, int], t=None) -> None:
        Y = self._reduce(None)["Y"]
        idx = [nmap.get(n, -1) for n in self.nodes]
        for a in range(len(idx)):
            for b in range(len(idx)):
                stamp_g(G, idx[a], idx[b], Y[a, b])

    def stamp_tran(self, G, I, x, x_prev, dt: float, nmap: Mapping[str, int], t=None) -> None:
        entry = self._reduce(dt)
        Y = entry["Y"]
        f = self._history(entry, dt)
        idx = [nmap.get(n, -1) for n in self.nodes]
        for a in range(len(idx)):
            for b in range(len(idx)):
                stamp_g(G, idx[a], idx[b], Y[a, b])
            stamp_i(I, idx[a], f[a])

    def ac_admittance(self, omega: float) -> np.ndarray:
        """The complex ``nE x nE`` admittance ``K_tt - K_ti K_ii^-1 K_it`` of
        ``K = L_sigma + j omega L_eps`` [S] (cached per ``omega``;
        ``fieldspice/circuit/coupling.py:216-230``).

        Column ``b`` is the unit excitation of electrode ``b``: the interior
        potential ``x + jy`` solves ``(B + jC)(x + jy) = -(B + jC) e_b`` on the
        interior (``presb_fgmres``, all columns together, every ``H`` solve one
        multi-RHS device solve), and ``Y[a, b]`` is the electrode-``a`` sum of
        ``K phi``.  ``Y(-omega) = conj(Y(omega))``; ``omega = 0`` is the DC
        stamp's admittance."""
        omega = float(omega)
        if omega == 0.0:
            return self._reduce(None)["Y"].astype(complex)
        if omega < 0.0:
            return np.conj(self.ac_admittance(-omega))
        if omega in self._ac_cache:
            return self._ac_cache[omega]
        nE = len(self.terminals)
        n = self.grid.n_nodes
        interior = (self._tid < 0).astype(np.float64)
        Ls, Le = self.L_sigma, self.L_eps
        sig = not self._sigma_zero
        H = Le.combine(omega, 1.0, Ls) if sig else Le.combine(omega)
        ls = GridLinearSystem(H, self._tid >= 0, self.cfg, grid=self.grid)

        def rows(op, X, scale=1.0):
            return np.stack([scale * op.apply(x) for x in X]) * interior

        def apply_B(X):
            return rows(Ls, X) if sig else np.zeros_like(X)

        def apply_C(X):
            return rows(Le, X, omega)

        def solve_H(X):
            return ls.solve(np.ascontiguousarray(X * interior)).reshape(X.shape)

        E = np.zeros((nE, n))                # unit excitations on the electrodes
        for k, t in enumerate(self.terminals):
            E[k, t.nodes] = 1.0
        x = np.zeros((nE, n))
        y = np.zeros((nE, n))
        if interior.any():
            x, y, info = presb_fgmres(apply_B, apply_C, solve_H, -apply_B(E), -apply_C(E),
                                      tol=self.cfg.tol)
            self.ac_info = dict(info, omega=omega)
        phr = E + x
        Y = np.zeros((nE, nE), dtype=complex)
        for b in range(nE):
            kr = -omega * Le.apply(y[b])         # K phi = (B + jC)(phr + j y)
            ki = omega * Le.apply(phr[b])
            if sig:
                kr += Ls.apply(phr[b])
                ki += Ls.apply(y[b])
            for a, t in enumerate(self.terminals):
                Y[a, b] = complex(kr[t.nodes].sum(), ki[t.nodes].sum())
        self._ac_cache[omega] = Y
        return Y

    def stamp_ac(self, Y_mat, J, x_op, omega: float, nmap: Mapping[str, int]) -> None:
        """Frequency domain: ``K = L_sigma + j omega L_eps`` needs no history
        (``fieldspice/circuit/coupling.py:216-237``)."""
        Y = self.ac_admittance(omega)
        idx = [nmap.get(n, -1) for n in self.nodes]
        for a in range(len(idx)):
            for b in range(len(idx)):
                stamp_g(Y_mat, idx[a], idx[b], Y[a, b])

    def accept_timestep(self, x, x_prev, dt: float, nmap: Mapping[str, int]) -> None:
        """Back-substitute the interior field of an accepted step
        (``fieldspice/circuit/coupling.py:239-252``)."""
        entry = self._reduce(dt)
        v = np.array([0.0 if nmap.get(n, -1) < 0 else float(x[nmap[n]]) for n in self.nodes])
        if entry["ls"] is not None and dt is not None:
            f = self.L_eps.apply(self.phi) / dt
            # the previous field starts the solve: a step changes it little
            self.phi = self._interior_solve(entry, f, v, guess=self.phi)
            self.last_iterations = int(entry["ls"].info.get("iterations", 0))
        else:
            phi = np.zeros(self.grid.n_nodes)
            phi[self._flat] = np.concatenate(
                [np.full(t.nodes.size, v[k]) for k, t in enumerate(self.terminals)])
            self.phi = phi

    def reset_state(self) -> None:
        """Per-Newton-iteration state: none.  The field history must survive
        (``fieldspice/circuit/coupling.py:254-268``)."""
        return None

    def reset_history(self) -> None:
        """Discard the stored field, starting a fresh transient."""
        self.phi = np.zeros(self.grid.n_nodes)

    def terminal_admittance(self, dt: float | None = None) -> np.ndarray:
        """The reduced ``nE x nE`` admittance [S] for a step ``dt`` (None: DC)."""
        return self._reduce(dt)["Y"]

    def op(self, x, nmap: Mapping[str, int]) -> dict[str, float]:
        return {f"{self.name}.v[{n}]": (0.0 if nmap.get(n, -1) < 0 else float(x[nmap[n]]))
                for n in self.nodes}
