"""Per-phase time of one iteration of the coarsest-level BiCGSTAB kernel (k_bottom) on the 3D
benchmark solve (octree of 8^3 blocks, root 4^3, levels 1-4 around an r = 0.2 sphere, 4 864
blocks; the level-1 problem is 64 blocks, 32 768 cells).

Runs the solves through lib/libsayaka_dev_phases.so, the device library built with the phase
timing compiled in (include/sayaka_phases.h; the product library has none of it), so it must
run in a process that has not loaded libsayaka_dev.so:

    python -m sayakaamr_amd.bottom_phases [--solves 5] [--out phases.json]

Prints, per precision (FP64: reference-order sums; mixed: the two-level preconditioned FP32
cycle's FP64 coarse solve) and placement (the launch on one XCD, the default, and spread over
the device: "_spread"), the microseconds per iteration of every phase (workgroup 0's wall clock,
100 MHz on MI355X) and per launch of the set-up.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json

import numpy as np

from . import _native as N

NAMES = {
    0: "p update + operand tile + v = A p + staging",
    16: "A: staging wait", 1: "A: ordered per-block chains", 2: "A: grid barrier", 3: "A: totals chain over blocks",
    4: "v ghosts + s, x update + operand tile + t = A s + staging",
    17: "B: staging wait", 5: "B: ordered per-block chains (2 sums)", 6: "B: grid barrier",
    7: "B: totals chain over blocks (2 sums)",
    8: "x, r update + staging",
    18: "C: staging wait", 9: "C: ordered per-block chains", 10: "C: grid barrier", 11: "C: totals chain over blocks",
    12: "r ghosts",
}
LAUNCH = {13: "first reduction: set-up + chains", 14: "first reduction: grid barrier",
          15: "first reduction: totals", 19: "first reduction: staging wait"}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--solves", type=int, default=5)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    N.DEV_LIB = N.DEV_PHASES_LIB   # before the first device context loads the library
    from . import device as D
    from . import workloads as W
    from .tree import AmrTree

    spec = W.C1_SPEC
    tb = AmrTree(3)
    tree = W.build_tree(tb, spec, W.shell_criterion())
    inp = W.c1_inputs(tree, spec, seed=0)
    dev = D.MacSolverGPU(3, spec.nb, tree["level"].shape[0] + 64, device=0)
    res = {}
    try:
        dev.prepare(tree, spec)
        dev.set_bc([W.WALL], [W.NEUMANN])
        dev.set_coefs(inp["vf"], inp["af"], inp["b"])
        dev.upload(D.RHS, inp["rhs"])
        fn = dev.lib.sayaka_dev_bottom_phases
        fn.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p]
        fn.restype = C.c_int
        out = np.zeros(32, np.uint64)
        khz = C.c_int(0)
        runs = [(name + ("" if one else "_spread"), prec, one) for one in (True, False)
                for name, prec in (("fp64", D.FP64), ("mixed", D.MIXED))]
        for name, prec, one in runs:
            dev.set_bottom_placement(one)
            dev.zero(D.SOL)
            dev.solve(precision=prec)   # warm-up
            if fn(0, out.ctypes.data, 1, C.byref(khz)) != 0:
                raise RuntimeError(dev.lib.sayaka_dev_last_error().decode())
            for _ in range(args.solves):
                dev.zero(D.SOL)
                dev.solve(precision=prec)
            if fn(0, out.ctypes.data, 1, C.byref(khz)) != 0:
                raise RuntimeError(dev.lib.sayaka_dev_last_error().decode())
            us = 1e3 / khz.value
            iters, launches = int(out[30]), int(out[31])
            per_it = {v: round(float(out[k]) * us / iters, 3) for k, v in NAMES.items()}
            per_launch = {v: round(float(out[k]) * us / launches, 3) for k, v in LAUNCH.items()}
            res[name] = dict(placement="one XCD" if one else "spread", bottom_placement=dev.bottom_placement(),
                             solves=args.solves, launches=launches, iterations=iters, wall_clock_khz=khz.value,
                             us_per_iteration=round(sum(per_it.values()), 3), per_iteration_us=per_it,
                             per_launch_us=per_launch)
            print(name, json.dumps(res[name], indent=1), flush=True)
    finally:
        dev.close()
        tb.close()
    if args.out:
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1)
    return res


if __name__ == "__main__":
    main()
