"""Shared pieces of the GPU time loops: Dirichlet groups and source bases."""

from __future__ import annotations

import numpy as np

from ..boundaries import WALLS, wall_nodes

__all__ = ["DirichletGroups", "source_basis", "device_time_loop"]


class DirichletGroups:
    """Fixed nodes as groups sharing one value per time point.

    Terminals and Dirichlet walls are claimed in the given order; a later
    claim wins a shared node, which reproduces fieldspice's "last entry wins"
    deduplication (``fieldspice/solvers/eqs.py:108-129``: walls after
    terminals; ``fieldspice/boundaries.py:963-1029``: later wall wins).
    """

    def __init__(self, grid, terminals=(), bc=None, walls_last: bool = True):
        self.grid = grid
        self.owner = np.full(grid.n_nodes, -1, dtype=np.int32)
        self.sources: list = []

        def claim_terms():
            for t in terminals:
                if t.driven != "voltage":
                    continue
                self.owner[t.nodes] = len(self.sources)
                self.sources.append(("terminal", t))

        def claim_walls():
            if bc is None:
                return
            for w in WALLS:
                wb = getattr(bc, w)
                if not wb.is_dirichlet_like:
                    continue
                nodes = wall_nodes(grid, w)
                if np.ndim(wb.value_at(0.0)) == 0:
                    self.owner[nodes] = len(self.sources)
                    self.sources.append(("wall", wb))
                else:
                    # a value per wall node: every node is a group of its own
                    first = len(self.sources)
                    self.owner[nodes] = first + np.arange(nodes.size, dtype=np.int32)
                    self.sources.extend([("wall_node", (wb, m, nodes.size))
                                         for m in range(nodes.size)])

        if walls_last:
            claim_terms()
            claim_walls()
        else:
            claim_walls()
            claim_terms()

    @property
    def ngroups(self) -> int:
        return len(self.sources)

    @property
    def fixed(self) -> np.ndarray:
        return np.flatnonzero(self.owner >= 0)

    def values_at(self, t: float) -> np.ndarray:
        out = np.zeros(len(self.sources))
        g = 0
        while g < len(self.sources):
            kind, src = self.sources[g]
            if kind == "terminal":
                out[g] = float(src.value_at(t))
                g += 1
            elif kind == "wall":
                v = np.asarray(src.value_at(t), dtype=float)
                if v.ndim != 0:
                    raise ValueError("a Dirichlet wall that was scalar at t = 0 returned an "
                                     f"array at t = {t}")
                out[g] = float(v)
                g += 1
            else:                                   # the nodes of an array-valued wall
                wb, _, count = src
                v = np.asarray(wb.value_at(t), dtype=float)
                try:
                    out[g:g + count] = np.broadcast_to(v, (count,))
                except ValueError as exc:
                    raise ValueError(f"Dirichlet value of shape {v.shape} cannot be broadcast "
                                     f"to the wall's {count} nodes") from exc
                g += count
        return out

    def node_values(self, t: float) -> np.ndarray:
        """(fixed node indices, their values) at time t."""
        vals = self.values_at(t)
        fixed = self.fixed
        return fixed, vals[self.owner[fixed]] if fixed.size else np.empty(0)


def source_basis(vectors, max_basis: int = 8):
    """Express a sequence of node vectors as ``scale[k] @ basis``.

    Vectors that are scalar multiples of an earlier basis vector reuse it, so a
    constant source or a source switched on and off costs one basis vector.
    Returns (basis (m, n), scales (len(vectors), m)).
    """
    basis: list[np.ndarray] = []
    norms: list[float] = []
    scales = np.zeros((len(vectors), max_basis))
    for k, v in enumerate(vectors):
        v = np.asarray(v, dtype=float).ravel()
        nv = float(np.linalg.norm(v))
        if nv == 0.0:
            continue
        placed = False
        for m, (b, nb) in enumerate(zip(basis, norms)):
            c = float(v @ b) / (nb * nb)
            if np.array_equal(v, c * b) or np.allclose(v, c * b, rtol=1e-15, atol=0.0):
                scales[k, m] = c
                placed = True
                break
        if not placed:
            if len(basis) == max_basis:
                raise NotImplementedError(
                    f"the source takes more than {max_basis} independent shapes over the run; "
                    "the GPU time loop supports sources that are combinations of a few fixed "
                    "node vectors")
            basis.append(v)
            norms.append(nv)
            scales[k, len(basis) - 1] = 1.0
    m = max(len(basis), 0)
    return (np.array(basis) if basis else np.zeros((0, 0))), scales[:, :m]


# values and sources handed to one device call are bounded: the values array
# and the source basis are copied to the device for the whole call
VALUES_BUDGET_BYTES = 512 << 20
MAX_SOURCE_BASIS = 8


def _fits(basis, v):
    """Index of the basis vector ``v`` is a multiple of, or -1."""
    nv = float(np.linalg.norm(v))
    for m, b in enumerate(basis):
        nb = float(np.linalg.norm(b))
        c = float(v @ b) / (nb * nb)
        if np.array_equal(v, c * b) or np.allclose(v, c * b, rtol=1e-15, atol=0.0):
            return m
    return -1 if nv > 0.0 else -2


def _chunk_ends(nsteps, ngroups, sources, store_every):
    """Last step of each device call.  A call ends where the next step's
    source would be the (MAX_SOURCE_BASIS + 1)-th independent shape, or its
    values would pass the budget; ends are moved back to multiples of
    ``store_every`` where that leaves a non-empty call."""
    per_step = max(1, 8 * ngroups)
    cap_values = max(1, VALUES_BUDGET_BYTES // per_step - 1)
    ends, k0 = [], 0
    while k0 < nsteps:
        basis, k = [], k0
        while k < nsteps and k - k0 < cap_values:
            v = None if sources is None else sources[k + 1]
            if v is not None and _fits(basis, v) == -1:
                if len(basis) == MAX_SOURCE_BASIS:
                    break
                basis.append(v)
            k += 1
        if k == k0:          # one step always fits
            k = k0 + 1
        if store_every > 1 and k < nsteps:
            aligned = (k // store_every) * store_every
            if aligned > k0:
                k = aligned
        ends.append(k)
        k0 = k
    return ends


def device_time_loop(sys, Aexp, x0, nsteps, *, group=None, values=None, sources=None,
                     src=None, src_scale=None, store_every=0, x_before=None, **kw):
    """``fs_transient_run`` over ``nsteps`` steps, in as many calls as the
    values and sources need (one in the common case).  ``values``: (nsteps+1,
    ngroups) Dirichlet group values; either ``sources`` (nsteps+1 node vectors,
    entry 0 unused, None entries zero) or a fixed basis ``src`` with per-step
    ``src_scale``.  Each later call continues from the last field (and the one
    before it, for extrapolated guesses); a projection guess rebuilds its basis
    in every call.  Outputs are stitched as one call would return them."""
    ng = 0 if values is None else int(np.asarray(values).shape[1])
    if sources is not None:
        ends = _chunk_ends(nsteps, ng, sources, store_every)
    else:
        ends = _chunk_ends(nsteps, ng, None, store_every)
    outs, k0, x, xb = [], 0, np.asarray(x0, dtype=float).ravel(), x_before
    n = x.size
    for k1 in ends:
        L = k1 - k0
        c_src, c_scale = src, None if src_scale is None else src_scale[k0:k1 + 1]
        if sources is not None:
            basis, rows = [], np.zeros((L + 1, MAX_SOURCE_BASIS))
            for j in range(1, L + 1):
                v = sources[k0 + j]
                if v is None:
                    continue
                v = np.asarray(v, dtype=float).ravel()
                m = _fits(basis, v)
                if m == -2:
                    continue
                if m == -1:
                    basis.append(v)
                    m = len(basis) - 1
                    rows[j, m] = 1.0
                else:
                    nb = float(np.linalg.norm(basis[m]))
                    rows[j, m] = float(v @ basis[m]) / (nb * nb)
            c_src = np.array(basis) if basis else None
            c_scale = rows[:, :len(basis)] if basis else None
        se = store_every
        if se > 1 and k0 % se != 0:
            se = 1                    # stored every step here, subsampled below
        out = sys.transient(Aexp, x, L, group=group,
                            values=None if values is None else values[k0:k1 + 1],
                            src=c_src, src_scale=c_scale, store_every=se, x_before=xb, **kw)
        out["_k0"], out["_se"] = k0, se
        outs.append(out)
        x, xb, k0 = out["x"], out["x_before"], k1
    if len(outs) == 1:
        return outs[0]
    merged = dict(outs[-1])
    for key in ("probe_i", "probe_v", "step_stats"):
        if outs[0][key] is None:
            continue
        merged[key] = np.concatenate([outs[0][key]] + [o[key][1:] for o in outs[1:]])
    if store_every > 0:
        rows = []
        for c, o in enumerate(outs):
            st = o["stored"]
            steps = o["_k0"] + o["_se"] * np.arange(st.shape[0])
            keep = (steps % store_every == 0) & ((steps > o["_k0"]) | (c == 0))
            rows.append(st[keep])
        merged["stored"] = np.concatenate(rows)
    stats = dict(outs[-1]["stats"])
    for key in ("iterations", "refreshes", "fallbacks", "pc_applies", "replacements"):
        stats[key] = sum(o["stats"][key] for o in outs)
    for key in ("residual", "residual_eq", "stop_reason"):
        stats[key] = max(o["stats"][key] for o in outs)
    stats["solve_time"] = sum(o["stats"]["solve_time"] for o in outs)
    stats["converged"] = int(all(o["stats"]["converged"] for o in outs))
    merged["stats"] = stats
    merged["rc"] = max(o["rc"] for o in outs)
    merged["calls"] = len(outs)
    return merged
