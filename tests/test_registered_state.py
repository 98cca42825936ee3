"""Registered states on the host tree (TreeStateData; sy_tree_state_*): the regrid's packing
moves their blocks on the host exactly as the reference's does (SayakaRegrid.cpp:530-591: a
surviving block's whole grown box to its new slot, new blocks and freed slots at 0).  The fill
of the new blocks runs on the GPU (tests/test_gpu_state_remap.py)."""
from __future__ import annotations

import numpy as np
import pytest

import golden_cases as GC
from sayakaamr_amd import workloads as W
from sayakaamr_amd.tree import AmrTree


def test_regrid_moves_registered_state_blocks(golden):
    g = golden("sremap2")
    spec = GC.remap_spec("sremap2")
    tb = AmrTree(2, max_blocks=4000)
    W.build_tree(tb, spec, GC.remap_criterion("sremap2", 0))
    idx = tb.state_register(None, 3, GC.FILL_NGROW, [0, 1, 2])
    before = tb.export()
    pre = GC.remap_input("sremap2", before, 0)
    tb.state_set_grown(idx, pre)
    assert np.array_equal(tb.state_get_grown(idx), pre)
    tb.set_refine_flags(g["c_step0_refine"])
    tb.set_coarsen_flags(g["c_step0_coarsen"])
    tb.regrid()
    origin = tb.regrid_origin()
    post = tb.state_get_grown(idx)
    keep = origin >= 0
    assert (~keep).any()
    assert np.array_equal(post[:, keep], pre[:, origin[keep]])
    assert not post[:, ~keep].any()


def test_registered_state_errors():
    spec = GC.remap_spec("sremap2")
    tb = AmrTree(2, max_blocks=4000)
    W.build_tree(tb, spec, GC.remap_criterion("sremap2", 0))
    idx = tb.state_register(0, 1, 2, [0])
    with pytest.raises(ValueError, match="state shape"):
        tb.state_set_grown(idx, np.zeros((1, 3, 1, 12, 12)))
    with pytest.raises(RuntimeError, match="face direction"):
        tb.state_register(2, 1, 2, [0])   # a 2D tree has no z faces
