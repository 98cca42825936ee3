"""Learned initial guess for the MAC solve: data generation, training and evaluation on the GPU.

The network and its kernels are in csrc/dev/cnn.h (C ABI: sayaka_cnn_*, include/sayaka_amd.h).
The workload follows the learned-warm-start configuration:
* trees: the 3D octree of the solve benchmark (8^3 blocks, root 4^3, levels 1..4) refined
  around a solid sphere with centre U([0.3, 0.7]^3) and radius U(0.1, 0.25);
* coefficients: density ratio 10^U(1, 3) (B = 1/rho, smoothed over 1.5 finest cells), cut-cell
  fractions;
* rhs: 8 Gaussian blobs (centres U([0, 1]^3), widths U(0.02, 0.1), amplitudes N(0, 1)) on the
  fluid, projected to zero mean.

Inputs are generated on the device (sayaka_gen_sphere_blobs).  Training targets are the
solutions of the plain solver (mixed precision to a 1e-10 relative residual), gauge-fixed to
zero volume-weighted mean over the fluid.  A guess is used only when its initial residual
(Linf) is below that of the zero guess; otherwise the solve starts from zero (the fallback).
Every solve, with or without the guess, runs to the same absolute target 1e-10 * |b|_inf.

    python -m sayakaamr_amd.warmstart train --n-train 2000 --n-test 200 --epochs 20 --out params.npz
    python -m sayakaamr_amd.warmstart study --n-test 200 --out study.json

`study` measures what any initial guess can gain in front of this solver, on the held-out
problems: the outer iterations from zero, from one V-cycle of the rhs (x1, the preconditioner's
own first step, about 1 % from the solution), from x1 plus exact parts of its remaining error
e1 = x* - x1 (the per-block means of e1; the per-block mean-free part of e1, which is what a
per-block network can see; 90 % of e1), and from the trained network's guess, forced and with
the fallback.  x* is the mixed solve taken to 1e-13 |b|_inf.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import device as D
from . import workloads as W
from .tree import AmrTree

SPEC = W.C1_SPEC
LEN2 = 2.5e-3          # length^2 scale of the target: phi / (max|rhs| * LEN2) is O(1)
NP = D.MacSolverGPU.CNN_NP
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PARAMS = os.path.join(HERE, "data", "cnn_c3.npz")


def log(msg):
    print(f"[warmstart] {msg}", flush=True)


def sample(rng):
    return dict(centre=rng.uniform(0.3, 0.7, 3), radius=float(rng.uniform(0.1, 0.25)),
                ratio=float(10.0 ** rng.uniform(1.0, 3.0)),
                blobs=dict(c=rng.uniform(0.0, 1.0, (8, 3)), w=rng.uniform(0.02, 0.1, 8), a=rng.normal(0.0, 1.0, 8)))


def init_params(seed=0):
    """He-initialised weights (used input / output channels only), zero biases."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NP, np.float32)
    pl = 27 * 16 * 16 + 16
    for l, (cin, cout) in enumerate(((5, 16), (16, 16), (16, 16), (16, 1))):
        w = np.zeros((27, 16, 16), np.float32)
        w[:, :cin, :cout] = rng.normal(0.0, np.sqrt(2.0 / (27 * cin)), (27, cin, cout))
        if l == 3:
            w *= 0.1
        p[l * pl:l * pl + 27 * 256] = w.ravel()
    return p


class Problem:
    """One random sphere / blob problem set up on a device context."""

    def __init__(self, dev: D.MacSolverGPU, s: dict, nsub=4):
        tb = AmrTree(3, max_blocks=dev.max_blocks)
        self.tree = tb.build_shell(SPEC, s["centre"], s["radius"], W.SHELL_FACTOR)
        dev.prepare(self.tree, SPEC)
        dev.set_bc([W.WALL], [W.NEUMANN])
        dxf = (SPEC.hi[0] - SPEC.lo[0]) / SPEC.ndiv[0] / SPEC.nb / 2 ** (SPEC.maxlevel - 1)
        dev.gen_sphere_blobs(self.tree, s["centre"], s["radius"], s["ratio"], 1.5 * dxf, nsub, s["blobs"])
        self.dev = dev
        self.L = int(self.tree["level"].max())
        self.leaves = np.nonzero(self.tree["node_type"] == 1)[0].astype(np.int32)
        _, self.bmax = dev.dot(self.L, D.RHS, D.RHS)        # |b|_inf over the finest tower

    def residual_inf(self, field):
        self.dev.residual(self.L, D.R, field, D.RHS)
        return self.dev.dot(self.L, D.R, D.R)[1]

    def gauge_shift(self, sol):
        """Volume-weighted mean of sol over the fluid leaves (the Neumann problem's constant)."""
        leaf = self.tree["node_type"] == 1
        dv = (self.tree["length"][:, 0] / SPEC.nb) ** 3
        w = (dv[:, None] * self.dev.get_vf())[leaf]
        return float(np.sum(w * sol[leaf]) / np.sum(w))


def make_device(max_blocks=12000, windows=4096):
    dev = D.MacSolverGPU(3, 8, max_blocks, device=0)
    dev.max_blocks = max_blocks
    dev.cnn_setup(windows)
    return dev


def generate(dev, seeds, per_solve, rng_pick):
    """Solve each problem, then store per_solve random leaf windows with their targets."""
    t0 = time.time()
    for q, seed in enumerate(seeds):
        pb = Problem(dev, sample(np.random.default_rng(seed)))
        dev.zero(D.SOL)
        st = dev.solve(tol_rel=1e-10, tol_abs=0.0, max_iter=200, precision=D.MIXED)
        if st["status"] != 0:
            log(f"seed {seed}: solve status {st['status']} (skipped)")
            continue
        sol = dev.download(D.SOL)
        shift = pb.gauge_shift(sol)
        pick = rng_pick.choice(pb.leaves, size=min(per_solve, pb.leaves.size), replace=False)
        n = dev.cnn_collect(pick, 1.0 / pb.bmax, shift, 1.0 / (pb.bmax * LEN2))
        if (q + 1) % 50 == 0 or q == len(seeds) - 1:
            log(f"generated {q + 1}/{len(seeds)} solves, {n} windows, {time.time() - t0:.0f} s")
    return n


def train(dev, nwin, epochs, batch, lr, seed=0):
    rng = np.random.default_rng(seed)
    hist = []
    t0 = time.time()
    for ep in range(epochs):
        perm = rng.permutation(nwin)
        losses = []
        for s in range(0, nwin - batch + 1, batch):
            losses.append(dev.cnn_train_step(perm[s:s + batch], lr))
        hist.append(float(np.mean(losses)))
        log(f"epoch {ep + 1}/{epochs}: mean loss {hist[-1]:.4e} ({time.time() - t0:.0f} s)")
    return hist


def solve_pair(dev, pb, precision):
    """Zero-guess solve and learned-guess solve (with fallback) to the same absolute target."""
    tol = 1e-10 * pb.bmax
    dev.zero(D.SOL)
    dev.sync()
    t0 = time.perf_counter()
    z = dev.solve(tol_rel=0.0, tol_abs=tol, max_iter=400, precision=precision)
    t_zero = time.perf_counter() - t0
    t0 = time.perf_counter()
    dev.cnn_guess(D.SOL, 1.0 / pb.bmax, pb.bmax * LEN2)
    r_guess = pb.residual_inf(D.SOL)
    used = r_guess < pb.bmax
    if not used:
        dev.zero(D.SOL)
    g = dev.solve(tol_rel=0.0, tol_abs=tol, max_iter=400, precision=precision)
    t_guess = time.perf_counter() - t0
    return dict(zero_iters=z["iters"], zero_ms=t_zero * 1e3, guess_iters=g["iters"], guess_ms=t_guess * 1e3,
                used=bool(used), r_guess_rel=float(r_guess / pb.bmax), status=(z["status"], g["status"]))


def evaluate(dev, seeds, precision=D.MIXED):
    out = []
    for seed in seeds:
        pb = Problem(dev, sample(np.random.default_rng(seed)))
        out.append(solve_pair(dev, pb, precision))
    return out


def study_problem(dev, pb):
    """Iterations (and initial residual / |b|_inf) of the mixed solve to 1e-10 |b|_inf from the
    guesses of the module docstring, for one problem set up on dev."""
    tol = 1e-10 * pb.bmax
    leaf = pb.tree["node_type"] == 1
    fl = leaf[:, None] & (dev.get_vf() > 0)

    def run(g):
        dev.upload(D.SOL, g)
        r0 = pb.residual_inf(D.SOL) / pb.bmax
        st = dev.solve(tol_rel=0.0, tol_abs=tol, max_iter=400, precision=D.MIXED)
        return dict(iters=int(st["iters"]), status=int(st["status"]), r0=float(r0))

    dev.zero(D.SOL)
    dev.solve(tol_rel=0.0, tol_abs=1e-3 * tol, max_iter=60, precision=D.MIXED)
    xs = dev.download(D.SOL)
    dev.zero(D.X)
    dev.precond(pb.L, D.X, D.RHS)
    x1 = dev.download(D.X)
    e1 = np.where(fl, xs - x1, 0.0)
    cnt = np.maximum(fl.sum(1), 1)
    bmean = np.where(fl, (e1.sum(1) / cnt)[:, None], 0.0)
    out = dict(zero=run(np.zeros_like(xs)), x1=run(x1))
    out["x1+block_means"] = run(x1 + bmean)
    out["x1+block_detail"] = run(x1 + (e1 - bmean))
    out["x1+0.9e1"] = run(x1 + 0.9 * e1)
    dev.cnn_guess(D.SOL, 1.0 / pb.bmax, pb.bmax * LEN2)
    g = dev.download(D.SOL)
    out["net"] = run(g)
    out["net_fallback"] = out["net"] if out["net"]["r0"] < 1.0 else dict(out["zero"])
    return out


def study(dev, seeds):
    rows = []
    for q, seed in enumerate(seeds):
        pb = Problem(dev, sample(np.random.default_rng(seed)))
        rows.append(study_problem(dev, pb))
        if (q + 1) % 20 == 0:
            log(f"study {q + 1}/{len(seeds)}")
    keys = list(rows[0])
    return dict(n=len(rows),
                mean_iters={k: float(np.mean([r[k]["iters"] for r in rows])) for k in keys},
                median_initial_resid_rel={k: float(np.median([r[k]["r0"] for r in rows])) for k in keys},
                all_converged=bool(all(r[k]["status"] == 0 for r in rows for k in keys)),
                net_used_frac=float(np.mean([r["net"]["r0"] < 1.0 for r in rows])))


def summarize(res):
    zi = np.array([r["zero_iters"] for r in res])
    gi = np.array([r["guess_iters"] for r in res])
    zt = np.array([r["zero_ms"] for r in res])
    gt = np.array([r["guess_ms"] for r in res])
    used = np.array([r["used"] for r in res])
    rr = np.array([r["r_guess_rel"] for r in res])
    return dict(n=len(res), zero_iters=float(zi.mean()), guess_iters=float(gi.mean()), zero_ms=float(np.median(zt)),
                guess_ms=float(np.median(gt)), used_frac=float(used.mean()),
                guess_resid_rel_median=float(np.median(rr)))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "eval", "study"])
    ap.add_argument("--n-train", type=int, default=2000)
    ap.add_argument("--n-test", type=int, default=200)
    ap.add_argument("--per-solve", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--params", default=DEFAULT_PARAMS)
    ap.add_argument("--out", default=None, help="trained parameters (train) / JSON record (eval, study)")
    args = ap.parse_args(argv)
    dev = make_device()
    test_seeds = list(range(1_000_000, 1_000_000 + args.n_test))
    if args.cmd == "study":
        dev.cnn_set_params(np.load(args.params)["params"])
        res = study(dev, test_seeds)
        log(f"study {json.dumps(res)}")
        if args.out:
            with open(args.out, "w") as f:
                json.dump(res, f, indent=1)
        dev.close()
        return res
    if args.cmd == "train":
        args.out = args.out or DEFAULT_PARAMS
        dev.cnn_set_params(init_params(0))
        dev.cnn_store_reset(args.n_train * args.per_solve)
        nwin = generate(dev, list(range(args.n_train)), args.per_solve, np.random.default_rng(1))
        hist = train(dev, nwin, args.epochs, args.batch, args.lr)
        params = dev.cnn_get_params()
        os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
        np.savez(args.out, params=params, loss_hist=np.array(hist), n_train=args.n_train,
                 per_solve=args.per_solve, epochs=args.epochs, lr=args.lr)
        log(f"saved {args.out}")
    else:
        dev.cnn_set_params(np.load(args.params)["params"])
    res = evaluate(dev, test_seeds)
    summ = summarize(res)
    log(f"held-out {summ}")
    if args.cmd == "eval" and args.out:
        with open(args.out, "w") as f:
            json.dump(summ, f, indent=1)
    dev.close()
    return summ


if __name__ == "__main__":
    main()
