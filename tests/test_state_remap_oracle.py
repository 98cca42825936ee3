"""Regrid remap of a registered state (AmrTree::regrid packing, SayakaRegrid.cpp:530-591, and
initNewBlockDataAfterRegrid, 987-1105): the oracle's restatement (oracle/fill_oracle.py
remap_state) against the reference's own output.

tests/golden/{s,f}remap{2,3}.npz: regrids of a moving ring (2D, 8^2 blocks) / shell (3D, 4^3
blocks) with a registered state -- cell data with 3 / 2 components (injection, centred linear,
limited; every wall its own code and BC mix) and face data of every direction on a periodic
tree (injection / centred linear / limited).  The fixture holds, per regrid, the whole grown
state of the blocks the remap writes (new blocks and parents of new blocks); every other block
holds its moved data, which the generator asserted against the reference's output.  The
product tree replays the same regrids (same flags) and gives each new block's origin.
"""
from __future__ import annotations

import numpy as np
import pytest

import golden_cases as GC
from oracle import fill_oracle as F
from sayakaamr_amd import workloads as W
from sayakaamr_amd.tree import AmrTree


def replay(name, golden, check):
    g = golden(name)
    spec = GC.remap_spec(name)
    dirs = GC.REMAP_CASES[name][3]
    for fd in (dirs if dirs is not None else (None,)):
        tag = "c" if fd is None else f"d{fd}"
        tb = AmrTree(spec.ndim, max_blocks=4000)
        W.build_tree(tb, spec, GC.remap_criterion(name, 0))
        for step in range(GC.REMAP_CASES[name][5]):
            before = tb.export()
            pre = GC.remap_input(name, before, step, fd)
            tb.set_refine_flags(g[f"{tag}_step{step}_refine"])
            tb.set_coarsen_flags(g[f"{tag}_step{step}_coarsen"])
            tb.regrid()
            after = tb.export()
            origin = tb.regrid_origin()
            touched = GC.remap_touched(after, origin)
            assert np.array_equal(np.nonzero(touched)[0], g[f"{tag}_step{step}_touched"])
            check(tb, spec, fd, step, pre, after, origin, touched, g[f"{tag}_step{step}_post"])


@pytest.mark.parametrize("name", ["sremap2", "sremap3", "fremap2", "fremap3"])
def test_remap_oracle_matches_reference(name, golden):
    def check(tb, spec, fd, step, pre, after, origin, touched, ref):
        post = F.remap_state(after, origin, pre, spec.nb, spec.ndim, GC.FILL_NGROW, spec.maxlevel,
                             GC.remap_prolong(name, fd), GC.remap_bcmaps(name) if fd is None else None, fd)
        assert np.array_equal(post[:, touched].view(np.int64), ref.view(np.int64)), (fd, step)
        keep = origin >= 0
        assert np.array_equal(post[:, keep & ~touched], pre[:, origin[keep & ~touched]])
    replay(name, golden, check)


@pytest.mark.parametrize("name,planted", [("sremap2", "parent_fill"), ("sremap3", "parent_fill"),
                                           ("fremap2", "parent_fill"), ("fremap2", "face_correction")])
def test_fixture_pins_each_step(name, planted, golden, monkeypatch):
    """Each step of the remap changes the reference's output: with the parents' in-level fill
    or the face correction left out, the restatement no longer matches some regrid."""
    if planted == "parent_fill":
        monkeypatch.setattr(F.StateFill, "fill_block", lambda self, *a, **k: None)
        monkeypatch.setattr(F.FaceFill, "fill_block", lambda self, *a, **k: None)
    else:
        monkeypatch.setattr(F.FaceFill, "correct_faces", lambda self, *a, **k: None)
    mism = []

    def check(tb, spec, fd, step, pre, after, origin, touched, ref):
        post = F.remap_state(after, origin, pre, spec.nb, spec.ndim, GC.FILL_NGROW, spec.maxlevel,
                             GC.remap_prolong(name, fd), GC.remap_bcmaps(name) if fd is None else None, fd)
        mism.append(not np.array_equal(post[:, touched].view(np.int64), ref.view(np.int64)))
    replay(name, golden, check)
    assert any(mism)
