"""Moving-boundary run on the device: a solid sphere translating through the domain, the tree
regridded around it every step, the previous step's pressure carried to the new tree and used
as the initial guess of the next solve.

Per step (the regrid / projection cycle of the reference's flow demo, Test7.h:411-418, 580-657):
  1. flag + regrid on the host tree (AmrTree::regrid, SayakaRegrid.cpp:26-976): leaves within
     factor * (block edge) of the moved shell are refined, the others coarsened, up to three
     passes per step; the block origins of the passes are composed;
  2. the new tables to the device (sayaka_dev_set_topology);
  3. the pressure to the new numbering (sayaka_dev_remap: survivors moved, new blocks injected
     from their parents -- the reference's packing and initNewBlockDataAfterRegrid);
  4. cut-cell coefficients of the moved sphere (density ratio 1000), generated on the device
     (sayaka_gen_sphere), and the rhs: by default the divergence of a seeded face velocity
     resampled on the new tree (sayaka_gen_velocity + sayaka_gen_divergence_rhs, walls at 0;
     the same seed every step, so the field only moves with the tree and the sphere), or
     Gaussian blobs (rhs="blobs", sayaka_gen_sphere_blobs);
  5. the mixed-precision solve to 1e-10 |b|_inf from the initial guess the policy picks, and,
     for comparison, from zero.  guess="carried" starts from the carried pressure, "zero" from
     zero, "auto" from the carried pressure only when its Linf residual is below the zero
     guess's (|b|_inf) -- the fallback rule of the learned-guess configuration, with the carried
     pressure as the candidate; the residual test is inside the timed solve;
  6. the judge: the true residual b - A x of the mixed solution under the FP64 operator (the
     reference's, bit for bit), and at the steps in fp64_steps the plain FP64 solve from zero
     beside it (its status, iterations, time and distance from the mixed solution);
  7. (projection rhs) the velocity correction u -= B flux(p) with coarse/fine faces
     re-synchronised (sayaka_dev_project), and the divergence left on the leaves.

    python -m sayakaamr_amd.moving --steps 10
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import device as D
from . import workloads as W
from .tree import AmrTree


def compose(first, second):
    """Origins of two regrids in a row: block b after the second came from block
    first[second[b]] before the first (-1: created by either)."""
    out = np.full(second.shape[0], -1, np.int32)
    m = second >= 0
    out[m] = first[second[m]]
    return out


def run(steps=10, spec=W.C4_SPEC, radius=W.C4_RADIUS, x0=0.5, shift=0.01, ratio=1000.0, compare_zero=True,
        max_blocks=45000, passes=3, rhs="projection", seed=1, log=None, fp64_steps=(), fp64_max_iter=100,
        guess="auto"):
    if rhs not in ("projection", "blobs"):
        raise ValueError(f"rhs must be 'projection' or 'blobs', not {rhs!r}")
    if guess not in ("carried", "zero", "auto"):
        raise ValueError(f"guess must be 'carried', 'zero' or 'auto', not {guess!r}")
    log = log or (lambda m: print(f"[moving] {m}", file=sys.stderr, flush=True))
    tb = AmrTree(spec.ndim, max_blocks=max_blocks)
    tree = tb.build_shell(spec, (x0, 0.5, 0.5), radius, W.SHELL_FACTOR)
    dev = D.MacSolverGPU(spec.ndim, spec.nb, max_blocks, device=0)
    rng = np.random.default_rng(7)
    blobs = dict(c=rng.uniform(0, 1, (8, 3)), w=rng.uniform(0.02, 0.1, 8), a=rng.normal(0, 1, 8))
    dxf = (spec.hi[0] - spec.lo[0]) / spec.ndiv[0] / spec.nb / 2 ** (spec.maxlevel - 1)
    rows = []
    try:
        for step in range(steps + 1):
            centre = (x0 + shift * step, 0.5, 0.5)
            row = dict(step=step)
            t0 = time.perf_counter()
            if step > 0:
                nold = tree["level"].shape[0]
                origin = np.arange(nold, dtype=np.int32)
                for _ in range(passes):
                    tb.flag_shell(centre, radius, W.SHELL_FACTOR)
                    n0 = tb.numblocks
                    tb.regrid()
                    o = tb.regrid_origin()
                    origin = compose(origin, o)
                    if tb.numblocks == n0 and np.array_equal(o, np.arange(n0)):
                        break
                tree = tb.export()
            t1 = time.perf_counter()
            dev.prepare(tree, spec)
            dev.set_bc([W.WALL], [W.NEUMANN])
            dev.sync()
            t2 = time.perf_counter()
            if step > 0:
                dev.remap(origin, nold, [D.SOL])
            t3 = time.perf_counter()
            if rhs == "blobs":
                dev.gen_sphere_blobs(tree, centre, radius, ratio, 1.5 * dxf, 4, blobs)
            else:
                dev.gen_sphere(tree, centre, radius, ratio, 1.5 * dxf, 4)
                dev.gen_velocity(tree, spec, seed, zero_walls=True)
                dev.gen_divergence_rhs()
            L = int(tree["level"].max())
            bmax = dev.dot(L, D.RHS, D.RHS)[1]
            dev.sync()
            t4 = time.perf_counter()
            tol = 1e-10 * bmax
            if step == 0:
                dev.zero(D.SOL)
            if compare_zero and step > 0:
                dev.copy(D.TMP, D.SOL)
                dev.zero(D.SOL)
                dev.sync()
                tz = time.perf_counter()
                rz = dev.solve(tol_rel=0.0, tol_abs=tol, max_iter=200, precision=D.MIXED)
                dev.sync()
                row.update(zero_ms=(time.perf_counter() - tz) * 1e3, zero_iters=rz["iters"])
                dev.copy(D.SOL, D.TMP)
            dev.sync()
            t5 = time.perf_counter()
            used = "zero"
            if step > 0 and guess != "zero":
                dev.residual(L, D.R, D.SOL, D.RHS)
                rc = dev.dot(L, D.R, D.R)[1]
                row["carried_resid_rel"] = float(rc / bmax)
                used = "carried" if guess == "carried" or rc < bmax else "zero"
            if used == "zero":
                dev.zero(D.SOL)
            row["guess"] = used
            r = dev.solve(tol_rel=0.0, tol_abs=tol, max_iter=200, precision=D.MIXED)
            dev.sync()
            t6 = time.perf_counter()
            # judge: the true residual of the mixed solution under the FP64 operator
            dev.residual(L, D.R, D.SOL, D.RHS)
            row["judge_rel_resid"] = float(dev.dot(L, D.R, D.R)[1] / bmax)
            if step in fp64_steps:
                sol_mixed = dev.download(D.SOL)
                dev.zero(D.SOL)
                dev.sync()
                tf = time.perf_counter()
                rf = dev.solve(tol_rel=0.0, tol_abs=tol, max_iter=fp64_max_iter, precision=D.FP64)
                dev.sync()
                row.update(fp64_ms=(time.perf_counter() - tf) * 1e3, fp64_iters=rf["iters"], fp64_status=rf["status"],
                           fp64_rel_resid=float(rf["rnorm"] / bmax))
                if rf["status"] == 0:
                    # fluid leaf cells only (a solid cell keeps whatever the initial guess held),
                    # each solution less its volume-weighted mean (the Neumann null space)
                    sol64 = dev.download(D.SOL)
                    vfl = dev.get_vf()
                    fluid = (tree["node_type"] == 1)[:, None] & (vfl > 0)
                    dv = np.broadcast_to(((tree["length"][:, 0] / spec.nb) ** spec.ndim)[:, None], vfl.shape)
                    wgt = (dv * vfl)[fluid]
                    a = sol_mixed[fluid] - np.sum(wgt * sol_mixed[fluid]) / np.sum(wgt)
                    b = sol64[fluid] - np.sum(wgt * sol64[fluid]) / np.sum(wgt)
                    row["mixed_rel_l2_vs_fp64"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
                dev.upload(D.SOL, sol_mixed)
            t7 = time.perf_counter()
            if rhs == "projection":
                div0, div1 = dev.project(D.SOL)
                row.update(project_ms=(time.perf_counter() - t7) * 1e3, div_before=div0, div_after=div1)
            leaf = tree["node_type"] == 1
            row.update(blocks=int(tree["level"].shape[0]), leaf_cells=int(leaf.sum()) * spec.nb ** spec.ndim,
                       regrid_ms=(t1 - t0) * 1e3, tables_ms=(t2 - t1) * 1e3, remap_ms=(t3 - t2) * 1e3,
                       coef_ms=(t4 - t3) * 1e3, solve_ms=(t6 - t5) * 1e3, iters=r["iters"], status=r["status"],
                       rel_resid=float(r["rnorm"] / bmax))
            rows.append(row)
            log(json.dumps({k: (float(f"{v:.4g}") if isinstance(v, float) else v) for k, v in row.items()}))
            if r["status"] != 0:
                raise RuntimeError(f"step {step}: solve status {r['status']}")
    finally:
        dev.close()
    return rows


def summarize(rows):
    s = [r for r in rows if r["step"] > 0]
    out = {k: float(np.mean([r[k] for r in s])) for k in ("regrid_ms", "tables_ms", "remap_ms", "coef_ms",
                                                            "solve_ms", "iters")}
    if s and "zero_ms" in s[0]:
        out["zero_guess_solve_ms"] = float(np.mean([r["zero_ms"] for r in s]))
        out["zero_guess_iters"] = float(np.mean([r["zero_iters"] for r in s]))
    out["step_ms"] = out["regrid_ms"] + out["tables_ms"] + out["remap_ms"] + out["coef_ms"] + out["solve_ms"]
    if s and "project_ms" in s[0]:
        out["project_ms"] = float(np.mean([r["project_ms"] for r in s]))
        out["step_ms"] += out["project_ms"]
        out["max_div_ratio"] = float(max(r["div_after"] / r["div_before"] for r in rows))
    out["steps"] = len(s)
    out["carried_used_frac"] = float(np.mean([r.get("guess") == "carried" for r in s])) if s else 0.0
    cr = [r["carried_resid_rel"] for r in s if "carried_resid_rel" in r]
    if cr:
        out["carried_resid_rel_median"] = float(np.median(cr))
    out["leaf_cells_mean"] = float(np.mean([r["leaf_cells"] for r in s]))
    out["max_rel_resid"] = float(max(r["rel_resid"] for r in s))
    out["max_judge_rel_resid"] = float(max(r["judge_rel_resid"] for r in rows))
    judged = [r for r in rows if "fp64_status" in r]
    if judged:
        out["fp64"] = [{k: r[k] for k in ("step", "fp64_status", "fp64_iters", "fp64_ms", "fp64_rel_resid",
                                          "mixed_rel_l2_vs_fp64") if k in r} for r in judged]
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--no-zero", action="store_true", help="skip the zero-guess comparison solve")
    ap.add_argument("--rhs", choices=("projection", "blobs"), default="projection")
    ap.add_argument("--fp64-steps", type=int, nargs="*", default=(), help="steps with the plain FP64 solve beside")
    ap.add_argument("--guess", choices=("carried", "zero", "auto"), default="auto")
    args = ap.parse_args(argv)
    rows = run(steps=args.steps, compare_zero=not args.no_zero, rhs=args.rhs, fp64_steps=tuple(args.fp64_steps),
               guess=args.guess)
    print(json.dumps(summarize(rows)))


if __name__ == "__main__":
    main()
