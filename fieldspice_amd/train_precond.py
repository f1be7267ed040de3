"""Train the learned preconditioner on the GPU.

    python -m fieldspice_amd.train_precond [--geoms 60] [--n 128] [--out PATH]

1. Seeded multi-conductor geometries (``workloads.multi_conductor``: uniform
   n^3 grid, 1-4 dielectric layers with eps_r in [2.5, 8], 2-8 conductor
   boxes).
2. Residual/error pairs from the FP64 solver.  For the unit excitation of
   each terminal, x* is solved to 1e-12, and CG runs stopped after k
   iterations give e_k = x* - x_k.  The library forms r = A e in FP64 and
   keeps the pairs on the device.
3. One shared network, trained round-robin over the geometries with Adam
   (``fs_pc_train``: forward cycle, energy-norm loss, backward through the
   cycle and the network, all on the GPU), with a cosine learning-rate decay.
4. Validation on unseen seeds: flexible-CG iterations and time with the
   network, with one Jacobi step at the bottom instead, and plain FP64 CG.

The trained parameters go to ``fieldspice_amd/weights/precond_net.npz``.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import _lib
from .precond import (DEFAULT_MAX_BOTTOM, DEFAULT_OMEGA, default_weights_path, hierarchy_groups,
                      save_net)
from .workloads import multi_conductor

DEFAULT_KS = (2, 5, 12, 30, 70, 160, 400)


def log(msg: str) -> None:
    print(f"[train {time.strftime('%H:%M:%S')}] {msg}", flush=True)


class Case:
    """One geometry on the device: operator, prepared system, terminals."""

    def __init__(self, ex):
        g = ex.grid
        self.ex = ex
        self.grid = g
        self.op = _lib.DeviceOperator.from_cells(g, ex.eps_cell)
        fixed = np.zeros(g.n_nodes, bool)
        for t in ex.terminals:
            fixed[t.nodes] = True
        self.fixed = fixed
        self.sys = _lib.DeviceSystem(self.op, fixed.astype(np.uint8))
        self.widths = (g.ax.hd, g.ay.hd, g.az.hd)

    def lifted(self, m: int) -> np.ndarray:
        """Right-hand side of terminal m at 1 V, the others at 0 V, after
        Dirichlet elimination."""
        xD = np.zeros(self.grid.n_nodes)
        xD[self.ex.terminals[m].nodes] = 1.0
        b = -self.op.apply(xD)
        b[self.fixed] = xD[self.fixed]
        return b

    def rhs(self) -> np.ndarray:
        return np.stack([self.lifted(m) for m in range(len(self.ex.terminals))])


def solver_pairs(case: Case, ks=DEFAULT_KS, tol=1e-12):
    """Exact errors of truncated FP64 CG runs (node space), (len(ks)*nt, n)."""
    B = case.rhs()
    xs, st, rc = case.sys.solve(B, _lib.DeviceSystem.params(tol))
    out = []
    for k in ks:
        xk, _, _ = case.sys.solve(B, _lib.DeviceSystem.params(tol, maxiter=k))
        out.append(xs - xk)
    return np.concatenate(out, axis=0), st


def solve_stats(case: Case, precision, pc=None, tol=1e-10, B=None):
    B = case.rhs() if B is None else B
    case.sys.set_precond(pc)
    t0 = time.perf_counter()
    x, st, rc = case.sys.solve(B, _lib.DeviceSystem.params(tol, precision=precision))
    wall = time.perf_counter() - t0
    return x, {"iterations": st["iterations"], "solve_time": st["solve_time"], "wall": wall,
               "residual_eq": st["residual_eq"], "fallbacks": st["fallbacks"], "rc": rc,
               "nrhs": B.shape[0]}


def evaluate(net, seeds, n, omega=DEFAULT_OMEGA, max_bottom=DEFAULT_MAX_BOTTOM, tol=1e-10):
    rows = []
    for s in seeds:
        case = Case(multi_conductor((n, n, n), seed=s))
        B = case.rhs()
        groups = hierarchy_groups(case.widths, max_bottom)
        pc_net = _lib.DevicePreconditioner(case.sys, net, groups, omega)
        pc_jac = _lib.DevicePreconditioner(case.sys, None, groups, omega)
        x64, r64 = solve_stats(case, _lib.FS_PREC_FP64, None, tol, B)
        _, rmx = solve_stats(case, _lib.FS_PREC_MIXED, None, tol, B)
        xj, rj = solve_stats(case, _lib.FS_PREC_FCG, pc_jac, tol, B)
        xl, rl = solve_stats(case, _lib.FS_PREC_FCG, pc_net, tol, B)
        err = float(np.abs(xl - x64).max() / max(np.abs(x64).max(), 1e-300))
        row = {"seed": s, "n": n, "nt": B.shape[0], "levels": len(groups),
               "fp64_cg": r64, "mixed_cg": rmx, "fcg_jacobi_bottom": rj, "fcg_net": rl,
               "err_vs_fp64": err}
        rows.append(row)
        log(f"eval n={n} seed={s} nt={B.shape[0]}: FP64 CG {r64['iterations']} it "
            f"{r64['solve_time']:.3f}s | mixed CG {rmx['iterations']} it {rmx['solve_time']:.3f}s | "
            f"FCG+Jacobi {rj['iterations']} it {rj['solve_time']:.3f}s | FCG+net "
            f"{rl['iterations']} it {rl['solve_time']:.3f}s  (speedup vs FP64 "
            f"{r64['solve_time'] / max(rl['solve_time'], 1e-12):.2f}x, err {err:.1e})")
        case.sys.set_precond(None)
        del pc_net, pc_jac
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--geoms", type=int, default=60)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--rounds", type=int, default=4)
    ap.add_argument("--steps", type=int, default=40, help="Adam steps per geometry per round")
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-min", type=float, default=1e-4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", default=None, help="start from these parameters (.npz)")
    ap.add_argument("--val-seeds", type=int, nargs="*", default=[100001, 100002])
    ap.add_argument("--val-n", type=int, nargs="*", default=[128])
    ap.add_argument("--out", default=default_weights_path())
    ap.add_argument("--max-pairs", type=int, default=2000)
    ap.add_argument("--max-bottom", type=int, default=DEFAULT_MAX_BOTTOM,
                    help="largest bottom grid (nodes) of the training hierarchies")
    a = ap.parse_args(argv)

    name, ncu = _lib.device_info()
    log(f"device {name} ({ncu} CUs); {a.geoms} geometries at {a.n}^3")
    net = _lib.DeviceNet(seed=a.seed)
    if a.init:
        with np.load(a.init, allow_pickle=False) as f:
            net.set_params(np.asarray(f["params"], np.float32))
    cases, pcs = [], []
    npairs = 0
    t0 = time.perf_counter()
    for gi in range(a.geoms):
        if npairs >= a.max_pairs:
            break
        case = Case(multi_conductor((a.n,) * 3, seed=a.seed * 100000 + gi))
        E, st = solver_pairs(case)
        E = E[:max(0, a.max_pairs - npairs)]
        groups = hierarchy_groups(case.widths, a.max_bottom)
        pc = _lib.DevicePreconditioner(case.sys, net, groups, DEFAULT_OMEGA)
        pc.add_pairs(E)
        npairs += E.shape[0]
        cases.append(case)
        pcs.append(pc)
        log(f"geometry {gi}: {len(case.ex.terminals)} terminals, FP64 CG {st['iterations']} it, "
            f"{E.shape[0]} pairs (total {npairs}), levels {len(groups)}, "
            f"{time.perf_counter() - t0:.1f}s")
    total = a.rounds * len(pcs)
    rng = np.random.default_rng(a.seed)
    hist = []
    step = 0
    t1 = time.perf_counter()
    for r in range(a.rounds):
        order = rng.permutation(len(pcs))
        round_losses = []
        for j, ci in enumerate(order):
            frac = step / max(total - 1, 1)
            lr = a.lr_min + 0.5 * (a.lr - a.lr_min) * (1 + math.cos(math.pi * frac))
            losses = pcs[ci].train(a.steps, a.batch, lr, seed=a.seed * 7919 + step)
            round_losses.append(float(losses.mean()))
            hist.append(float(losses.mean()))
            step += 1
            if j % 10 == 0:
                log(f"round {r} geometry {ci}: lr {lr:.2e} loss {losses.mean():.4f}")
        log(f"round {r} done: mean loss {np.mean(round_losses):.4f} "
            f"({time.perf_counter() - t1:.1f}s)")
    train_time = time.perf_counter() - t1
    # free the training systems before validation
    for pc in pcs:
        pc.clear_pairs()
    del pcs, cases
    save_net(net, a.out, n=a.n, geoms=step // max(a.rounds, 1), pairs=npairs,
             rounds=a.rounds, steps=a.steps, batch=a.batch, lr=a.lr,
             loss_history=np.asarray(hist), train_time=train_time)
    log(f"saved {a.out} ({net.num_params} parameters) after {train_time:.1f}s of training")
    results = {}
    for n in a.val_n:
        results[str(n)] = evaluate(net, a.val_seeds, n, max_bottom=a.max_bottom)
    summary = os.path.splitext(a.out)[0] + "_validation.json"
    with open(summary, "w") as f:
        json.dump({"train": {"pairs": npairs, "geometries": len(hist) // max(a.rounds, 1),
                             "n": a.n, "loss_first": hist[:5], "loss_last": hist[-5:],
                             "train_time": train_time},
                   "validation": results}, f, indent=1)
    log(f"validation summary in {summary}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
