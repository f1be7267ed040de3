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
"""

from __future__ import annotations

from typing import Any, Mapping, Sequence

import numpy as np

from .. import _lib
from ..solvers.base import SolverConfig, Terminal
from ..solvers.linear import GridLinearSystem

__all__ = ["FieldRegion", "stamp_g", "stamp_i"]


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
        ls._judge(st, rc)
        self.asymmetry = float(np.abs(Y - Y.T).max() / max(np.abs(Y).max(), 1e-300))
        entry = {"K": K, "Y": 0.5 * (Y + Y.T), "ls": ls, "dt": dt}
        self._cache[key] = entry
        return entry

    def _interior_solve(self, entry, f, v) -> np.ndarray:
        """``phi`` with the electrodes at ``v`` and ``K_ii phi_i = f_i - K_it v``."""
        K, ls = entry["K"], entry["ls"]
        x0 = np.zeros(self.grid.n_nodes)
        vals = np.concatenate([np.full(t.nodes.size, v[k]) for k, t in enumerate(self.terminals)])
        x0[self._flat] = vals
        b = f - K.apply(x0)
        b[self._flat] = vals
        return ls.solve(b)

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

    def stamp_ac(self, Y_mat, J, x_op, omega: float, nmap: Mapping[str, int]) -> None:
        raise NotImplementedError(
            "the AC stamp needs complex K = L_sigma + j omega L_eps; the GPU path solves real "
            "systems only")

    def accept_timestep(self, x, x_prev, dt: float, nmap: Mapping[str, int]) -> None:
        """Back-substitute the interior field of an accepted step
        (``fieldspice/circuit/coupling.py:239-252``)."""
        entry = self._reduce(dt)
        v = np.array([0.0 if nmap.get(n, -1) < 0 else float(x[nmap[n]]) for n in self.nodes])
        if entry["ls"] is not None and dt is not None:
            f = self.L_eps.apply(self.phi) / dt
            self.phi = self._interior_solve(entry, f, v)
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
