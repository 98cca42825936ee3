"""Benchmark of the MAC Poisson solve on one MI355X.

Workload (one "step" = one solve): the 3D variable-coefficient octree solve -- unit cube,
root 4^3 blocks of 8^3 cells, refined to tree level 4 around a sphere of radius 0.2 centred
at (0.5, 0.5, 0.5) (4 264 leaf blocks, 2.18 M leaf cells), density 1 outside / 1000 inside
through a smoothed Heaviside of width 1.5 finest dx, cut-cell volume / area fractions of
the sphere, rhs = divergence of a seeded U(-1,1) face velocity projected to zero mean,
Neumann walls, MG-preconditioned BiCGSTAB from a zero guess to 1e-10 relative residual.
Inputs are resident in HBM before the timed region; each step zeroes the solution on the
device and solves.

Reported: plain FP64 ms/solve (the headline value), the mixed-precision ms/solve over 5
seeded right-hand sides with their error against the FP64 solution, per-kernel-class
times with achieved algorithmic GB/s, the roofline of the dominant kernel class measured
with HIP events on the solver's stream inside the timed region, and the CPU baseline (the
compiled reference solver, or the C oracle when the reference build is absent, 1 core).

    python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload():
    from sayakaamr_amd import workloads as W
    from sayakaamr_amd.tree import AmrTree

    spec = W.C1_SPEC
    t = AmrTree(3)
    tree = W.build_tree(t, spec, W.shell_criterion())
    inp = W.c1_inputs(tree, spec, seed=0)
    return spec, tree, inp


def gauge(sol, tree, spec):
    """Remove the Neumann null space: subtract the value of the pinned cell."""
    from sayakaamr_amd import workloads as W

    b, c = W.pin_cell(tree, spec)
    return sol - sol[b, c]


def rel_l2(a, b, leaf):
    return float(np.linalg.norm(a[leaf] - b[leaf]) / np.linalg.norm(b[leaf]))


def time_solves(dev, D, precision, steps, prof_class=None):
    if prof_class is not None:
        dev.profile(prof_class, True)
    dev.sync()
    t0 = time.perf_counter()
    res = None
    for _ in range(steps):
        dev.zero(D.SOL)
        res = dev.solve(tol_rel=1e-10, tol_abs=0.0, max_iter=2000, precond=1, precision=precision)
        if res["status"] != 0:
            raise RuntimeError(f"solve did not converge: {res}")
    dev.sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    prof = dev.profile_read(prof_class) if prof_class is not None else None
    if prof_class is not None:
        dev.profile(prof_class, False)
    return ms, res, prof


def class_profile(dev, D, precision):
    """One solve with every kernel class timed (events around each launch)."""
    for k in range(len(D.KCLASS_NAMES)):
        dev.profile(k, True)
    dev.zero(D.SOL)
    res = dev.solve(tol_rel=1e-10, tol_abs=0.0, max_iter=2000, precond=1, precision=precision)
    out = {}
    for k, name in enumerate(D.KCLASS_NAMES):
        p = dev.profile_read(k)
        dev.profile(k, False)
        if p["launches"] == 0:
            continue
        gbs = p["alg_bytes"] / (p["total_ms"] * 1e-3) / 1e9 if p["alg_bytes"] > 0 and p["total_ms"] > 0 else None
        out[name] = dict(ms=round(p["total_ms"], 3), launches=p["launches"],
                         alg_gbs=None if gbs is None else round(gbs, 1))
    return res, out


def cpu_baseline(spec, tree, inp, want_solution=True):
    """The reference's FP64 cg_solve on one core (bounded: one C1 solve)."""
    from oracle import refbind as R
    from sayakaamr_amd import workloads as W

    n = tree["level"].shape[0]
    if R.available(3):
        rt = R.RefTree(3)
        rtree = W.build_tree(rt, spec, W.shell_criterion())
        if not np.array_equal(rtree["level"], tree["level"]) or not np.array_equal(rtree["neighbor"], tree["neighbor"]):
            raise RuntimeError("reference tree differs from the product tree")
        mac = R.RefMac(rt)
        mac.prepare()
        mac.set_ib_frac(inp["vf"], inp["af"])
        mac.set_coefb(inp["b"])
        mac.set_bc([W.WALL], [W.NEUMANN], [0.0])
        mac.set_rhs(inp["rhs"])
        mac.set_zero_sol()
        mac.set_params(1e-10, 0.0, 2000, 1)
        t0 = time.perf_counter()
        res = mac.cg_solve()
        ms = (time.perf_counter() - t0) * 1e3
        sol = mac.get_sol(n, spec.nb ** 3)
        kind = "reference"
    else:
        from oracle import oraclebind as O

        orc = O.Oracle(tree, spec)
        orc.prepare()
        orc.set_ib_frac(inp["vf"], inp["af"])
        orc.set_coefb(inp["b"])
        orc.set_bc([W.WALL], [W.NEUMANN], [0.0])
        orc.set_rhs(inp["rhs"])
        orc.set_zero_sol()
        orc.set_params(1e-10, 0.0, 2000, 1)
        t0 = time.perf_counter()
        res = orc.cg_solve()
        ms = (time.perf_counter() - t0) * 1e3
        sol = orc.get_sol()
        kind = "port"
    return ms, res, sol, kind


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seeds", type=int, default=5, help="right-hand sides of the mixed-precision leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); the solve is "
              "measured at N=1 and its multi-GPU sharding is a design note, not an implemented path.",
              file=sys.stderr)
        sys.exit(2)

    from sayakaamr_amd import device as D
    from sayakaamr_amd import workloads as W

    t0 = time.perf_counter()
    spec, tree, inp = build_workload()
    n = tree["level"].shape[0]
    leaf = tree["node_type"] == 1
    nleaf_cells = int(leaf.sum()) * spec.nb ** 3
    log(f"workload: {n} blocks, {int(leaf.sum())} leaves, {nleaf_cells} leaf cells ({time.perf_counter() - t0:.1f} s)")

    dev = D.MacSolverGPU(3, spec.nb, n + 64, device=0)
    dev.prepare(tree, spec)
    dev.set_bc([W.WALL], [W.NEUMANN])
    dev.set_coefs(inp["vf"], inp["af"], inp["b"])
    dev.upload(D.RHS, inp["rhs"])

    # warmup + per-class profile (every launch bracketed by events)
    for _ in range(args.warmup):
        dev.zero(D.SOL)
        dev.solve(precision=D.FP64)
    res64, classes64 = class_profile(dev, D, D.FP64)
    log(f"fp64 classes: {classes64}")
    timed = {k: v["ms"] for k, v in classes64.items()}
    dom_name = max(timed, key=timed.get)
    dom = D.KCLASS_NAMES.index(dom_name)
    hbm_classes = {k: v for k, v in timed.items() if classes64[k]["alg_gbs"] is not None}
    roof_name = dom_name if classes64[dom_name]["alg_gbs"] is not None else max(hbm_classes, key=hbm_classes.get)
    roof = D.KCLASS_NAMES.index(roof_name)

    # timed region: FP64 solves, events only around the roofline class's launches
    ms64, r64, prof = time_solves(dev, D, D.FP64, args.steps, prof_class=roof)
    sol64 = dev.download(D.SOL)
    log(f"fp64: {ms64:.2f} ms/solve, iters {r64['iters']}, rnorm {r64['rnorm']:.3e} / rnorm0 {r64['rnorm0']:.3e}")
    achieved = prof["alg_bytes"] / (prof["total_ms"] * 1e-3) / 1e9
    avg_launch_us = prof["total_ms"] * 1e3 / prof["launches"]
    bytes_per_launch = prof["alg_bytes"] / prof["launches"]

    # mixed precision: per seed, FP64 GPU solve (bitwise the reference's) vs mixed solve
    mixed = []
    ms_mixed_all = []
    for seed in range(args.seeds):
        rhs = inp["rhs"] if seed == 0 else W.c1_rhs(tree, spec, seed, inp["vf"], inp["af"])
        dev.upload(D.RHS, rhs)
        dev.zero(D.SOL)
        rf = dev.solve(precision=D.FP64)
        ref = gauge(dev.download(D.SOL), tree, spec)
        dev.zero(D.SOL)
        dev.solve(precision=D.MIXED)  # warm
        msm, rm, _ = time_solves(dev, D, D.MIXED, args.steps)
        solm = gauge(dev.download(D.SOL), tree, spec)
        ms_mixed_all.append(msm)
        mixed.append(dict(seed=seed, ms=round(msm, 3), iters=rm["iters"], fp64_iters=rf["iters"],
                          rel_resid=float(rm["rnorm"] / rm["rnorm0"]),
                          rel_l2_vs_fp64=rel_l2(solm, ref, leaf)))
        log(f"mixed seed {seed}: {msm:.2f} ms, iters {rm['iters']} (fp64 {rf['iters']}), "
            f"rel L2 vs fp64 {mixed[-1]['rel_l2_vs_fp64']:.2e}")
    dev.upload(D.RHS, inp["rhs"])
    _, classes_mixed = class_profile(dev, D, D.MIXED)
    dev.close()

    # V-cycle rate: preconditioner time / V-cycles of the FP64 solve
    nvc = 2 * r64["iters"] - (1 if r64["half_step"] else 0)
    vc_ms = sum(classes64.get(k, {}).get("ms", 0.0) for k in ("smooth", "bottom", "transfer"))
    cpu = None
    err_cpu = None
    if not args.no_cpu:
        log("cpu baseline: one FP64 solve on 1 core ...")
        cms, cres, csol, kind = cpu_baseline(spec, tree, inp)
        err_cpu = rel_l2(gauge(sol64, tree, spec), gauge(csol, tree, spec), leaf)
        cpu = dict(value=round(cms, 1), unit="ms/solve", cores=1, kind=kind,
                   sample=f"one full solve of the same workload (iters {cres['iters']}), "
                          f"{cpu_model()}, {os.cpu_count()} cpus visible, 1 used")
        log(f"cpu {kind}: {cms:.0f} ms/solve, iters {cres['iters']}; fp64 gpu rel L2 vs cpu {err_cpu:.2e}")

    out = {
        "metric": "ms/solve to 1e-10 residual, FP64 vs mixed+learned MAC Poisson; HBM GB/s",
        "value": round(ms64, 3),
        "unit": "ms/solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms64, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "3D octree MAC Poisson, 8^3 blocks, root 4^3, levels 1-4 around a r=0.2 sphere, "
                               "density ratio 1000, cut-cell fractions, Neumann walls, zero guess, tol 1e-10",
                   "blocks": n, "leaf_blocks": int(leaf.sum()), "leaf_cells": nleaf_cells,
                   "parallelism": "single GPU"},
        "fp64": {"iters": r64["iters"], "rel_resid": float(r64["rnorm"] / r64["rnorm0"]),
                 "bottom_calls": r64["bottom_calls"], "bottom_iters": r64["bottom_iters"],
                 "rel_l2_vs_cpu": err_cpu, "vcycles": nvc,
                 "vcycle_ms": round(vc_ms / max(nvc, 1), 4),
                 "vcycle_leaf_cells_per_s": round(nleaf_cells / (vc_ms / max(nvc, 1) * 1e-3), 1),
                 "kernels": classes64},
        "mixed": {"ms_per_solve": round(float(np.mean(ms_mixed_all)), 3),
                  "speedup_vs_fp64": round(ms64 / float(np.mean(ms_mixed_all)), 3),
                  "dtype": "f32 V-cycle, f64 outer", "seeds": mixed, "kernels": classes_mixed},
        "roofline": {"bound": "hbm", "kernel": roof_name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "launches": prof["launches"], "avg_launch_us": round(avg_launch_us, 3),
                     "alg_bytes_per_launch": round(bytes_per_launch, 1),
                     "dominant_class": dom_name},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
