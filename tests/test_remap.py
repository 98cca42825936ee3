"""Regrid data movement: the product tree's record of where every block came from
(AmrTree::regridOrigin, sy_tree_regrid_origin) and the oracle's restatement of the reference's
packing + initNewBlockDataAfterRegrid, against the reference's own output for a moving sphere
(3D, 4^3 blocks) and a moving ring (2D, 8^2 blocks), 6 regrids each with refinement and
coarsening (tests/golden/regrid_remap{2,3}.npz).  Bit-exact: every value is a distinct cell tag,
so a wrong origin, child offset or level order shows up as a wrong tag."""
from __future__ import annotations

import numpy as np
import pytest

from golden_cases import TREE_KEYS
from oracle import remap_oracle as RO
from sayakaamr_amd import workloads as W
from sayakaamr_amd.tree import AmrTree


def replay(ndim, golden):
    g = golden(f"regrid_remap{ndim}")
    spec = W.small_spec(ndim, 4, nb=4 if ndim == 3 else 8)
    if ndim == 3:
        crit0 = W.shell_criterion(centre=(0.35, 0.5, 0.5), r0=0.18)
    else:
        crit0 = W.ring_criterion(centre=(0.35, 0.5), r0=0.2, half=0.04)
    tb = AmrTree(ndim, max_blocks=4000)
    W.build_tree(tb, spec, crit0)
    for step in range(6):
        tb.set_refine_flags(g[f"step{step}_refine"])
        tb.set_coarsen_flags(g[f"step{step}_coarsen"])
        tb.regrid()
        yield step, spec, g, tb.export(), tb.regrid_origin()


@pytest.mark.parametrize("ndim", [3, 2])
def test_tree_after_each_regrid_matches_reference(ndim, golden):
    for step, spec, g, t, origin in replay(ndim, golden):
        for k in TREE_KEYS:
            if f"step{step}_{k}" in g:
                assert np.array_equal(t[k], g[f"step{step}_{k}"]), (step, k)


@pytest.mark.parametrize("ndim", [3, 2])
def test_origin_and_remap_match_reference(ndim, golden):
    for step, spec, g, t, origin in replay(ndim, golden):
        pre, post = g[f"step{step}_pre"], g[f"step{step}_post"]
        assert origin.shape[0] == post.shape[0]
        assert np.all(origin < pre.shape[0])
        kept = origin[origin >= 0]
        assert np.unique(kept).size == kept.size               # a block lands in one place
        assert np.all(np.diff(kept) > 0)                       # survivors keep their order
        got = RO.remap(pre, origin, t, spec.nb, ndim)
        assert np.array_equal(got, post), step


def test_injection_closed_form():
    """A parent whose cells hold their own centre coordinate gives each child cell the
    coordinate of the parent cell containing it: floor((fine index + offset) / 2)."""
    nb = 8
    for ndim in (2, 3):
        nz = nb if ndim == 3 else 1
        parent = np.arange(nb * nb * nz, dtype=np.float64)
        for wc in range(2 ** ndim):
            f = RO.inject(parent, wc, nb, ndim).reshape(nz, nb, nb)
            h = nb // 2
            for k in range(nz):
                for j in range(nb):
                    for i in range(nb):
                        ic = (i + (wc & 1) * nb) // 2
                        jc = (j + ((wc >> 1) & 1) * nb) // 2
                        kc = (k + ((wc >> 2) & 1) * nb) // 2 if ndim == 3 else 0
                        assert f[k, j, i] == parent[(kc * nb + jc) * nb + ic], (ndim, wc, i, j, k, h)


def test_compose_origins():
    from sayakaamr_amd import moving as MV
    first = np.array([0, 1, -1, 3, 4], np.int32)      # block 2 new in the first regrid
    second = np.array([0, 2, 3, -1, 4], np.int32)     # block 3 new in the second
    assert np.array_equal(MV.compose(first, second), [0, -1, 3, -1, 4])
