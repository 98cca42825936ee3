"""Regrid data movement on the GPU (sayaka_dev_remap) against the reference's own output: the
registered cell state after regrid + initNewBlockDataAfterRegrid for a moving sphere (3D, 4^3
blocks) and ring (2D, 8^2 blocks), 6 regrids each (tests/golden/regrid_remap{2,3}.npz).
Bit-exact (every value is a distinct cell tag)."""
from __future__ import annotations

import numpy as np
import pytest

from sayakaamr_amd import device as D
from sayakaamr_amd import workloads as W
from sayakaamr_amd.tree import AmrTree

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ndim", [3, 2])
def test_device_remap_matches_reference(ndim, golden):
    g = golden(f"regrid_remap{ndim}")
    spec = W.small_spec(ndim, 4, nb=4 if ndim == 3 else 8)
    if ndim == 3:
        crit0 = W.shell_criterion(centre=(0.35, 0.5, 0.5), r0=0.18)
    else:
        crit0 = W.ring_criterion(centre=(0.35, 0.5), r0=0.2, half=0.04)
    tb = AmrTree(ndim, max_blocks=4000)
    W.build_tree(tb, spec, crit0)
    print(f"[gpu] remap {ndim}D: 6 regrids", flush=True)
    dev = D.MacSolverGPU(ndim, spec.nb, 4000, device=0)
    try:
        for step in range(6):
            before = tb.export()
            pre = g[f"step{step}_pre"]
            assert pre.shape[0] == before["level"].shape[0]
            dev.prepare(before, spec)
            dev.upload(D.SOL, pre)
            dev.upload(D.RHS, -pre)
            tb.set_refine_flags(g[f"step{step}_refine"])
            tb.set_coarsen_flags(g[f"step{step}_coarsen"])
            tb.regrid()
            after = tb.export()
            dev.prepare(after, spec)
            dev.remap(tb.regrid_origin(), pre.shape[0], [D.SOL, D.RHS])
            post = g[f"step{step}_post"]
            assert np.array_equal(dev.download(D.SOL), post), step
            assert np.array_equal(dev.download(D.RHS), -post), step
    finally:
        dev.close()
