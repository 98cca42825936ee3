"""A registered state kept on the device (TreeData::setResident, no reference counterpart):
AmrTree::regrid's block moves, initNewBlockDataAfterRegrid and FillPatch::fillBoundary
(SayakaRegrid.cpp:530-591, 987-1105, SayakaFillPatch.cpp:17-368) run on the device copy.

* The reference's recorded regrids (tests/golden/{s,f}remap*.npz) replayed with the state
  resident: bit-exact on the blocks the remap writes, moved data everywhere else -- the same
  assertions as the host path's test.
* A per-step loop in the manner of the reference's flow demo (Test7.h:580-657: regrid around a
  moving shell, new blocks filled, ghost fill every step) on a 3D 8^3-block tree: the resident
  state ends bit-identical to the same loop on host data, moves no state bytes between the
  initial upload and the final download, and the tree's tables go up once per changed grid
  (not once per fill).
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

import golden_cases as GC
from sayakaamr_amd import workloads as W
from sayakaamr_amd.tree import AmrTree

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["sremap2", "sremap3", "fremap2"])
def test_resident_registered_state_matches_reference(name, golden):
    g = golden(name)
    spec = GC.remap_spec(name)
    dirs = GC.REMAP_CASES[name][3]
    for fd in (dirs if dirs is not None else (None,)):
        tag = "c" if fd is None else f"d{fd}"
        tb = AmrTree(spec.ndim, max_blocks=4000)
        try:
            W.build_tree(tb, spec, GC.remap_criterion(name, 0))
            prolong = GC.remap_prolong(name, fd)
            if fd is None:
                maps = GC.remap_bcmaps(name)
                codes = list(GC.FILL_WALLS)
                idx = tb.state_register(None, len(prolong), GC.FILL_NGROW, prolong, codes,
                                        [maps[c][k][0] for c in range(len(prolong)) for k in codes],
                                        [maps[c][k][1] for c in range(len(prolong)) for k in codes])
            else:
                idx = tb.state_register(fd, len(prolong), GC.FILL_NGROW, prolong)
            tb.state_set_resident(idx, True)
            for step in range(GC.REMAP_CASES[name][5]):
                before = tb.export()
                pre = GC.remap_input(name, before, step, fd)
                tb.state_set_grown(idx, pre)            # host -> the device copy
                tb.set_refine_flags(g[f"{tag}_step{step}_refine"])
                tb.set_coarsen_flags(g[f"{tag}_step{step}_coarsen"])
                tb.regrid()
                print(f"[gpu] resident registered state {name} {tag} step {step}", flush=True)
                tb.init_new_data()
                after = tb.export()
                origin = tb.regrid_origin()
                post = tb.state_get_grown(idx)          # the device copy -> host
                touched = GC.remap_touched(after, origin)
                assert np.array_equal(post[:, touched].view(np.int64),
                                      g[f"{tag}_step{step}_post"].view(np.int64)), (fd, step)
                keep = ~touched
                assert np.array_equal(post[:, keep], pre[:, origin[keep]])
        finally:
            tb.close()


LOOP_SPEC = dataclasses.replace(W.small_spec(3, maxlevel=4, nb=8), bc=GC.FILL_WALLS, ngrow=2)
LOOP_PROLONG = (1, 2)   # centred linear, MC-limited centred linear


def fill_loop(resident, steps=6):
    """Regrid around a shell moving +0.03 in x, new blocks filled, a 2-layer ghost fill every
    step; returns the final grown state, the traffic of the loop and the per-step block counts."""
    tb = AmrTree(3, max_blocks=8000)
    try:
        centre = np.array([0.40, 0.5, 0.5])
        tb.build_shell(LOOP_SPEC, centre, 0.2, W.SHELL_FACTOR)
        maps = GC.remap_bcmaps("sremap3")
        codes = list(GC.FILL_WALLS)
        idx = tb.state_register(None, 2, 2, LOOP_PROLONG, codes,
                                [maps[c][k][0] for c in range(2) for k in codes],
                                [maps[c][k][1] for c in range(2) for k in codes])
        if resident:
            tb.state_set_resident(idx, True)
        shp = tb._state_shape(idx)
        rng = np.random.default_rng(7)
        x0 = np.zeros(shp)
        x0[:, :, 2:-2, 2:-2, 2:-2] = rng.uniform(-1.0, 1.0, (shp[0], shp[1], 8, 8, 8))
        tb.state_set_grown(idx, x0)
        tb.state_fill(idx, 2)
        s0 = tb.transfer_stats()
        nblk = []
        for step in range(steps):
            centre[0] += 0.03
            tb.flag_shell(centre, 0.2, W.SHELL_FACTOR)
            tb.regrid()
            tb.init_new_data()
            tb.state_fill(idx, 2)
            nblk.append(tb.numblocks)
        tb.state_fill(idx, 2)        # a fill with no regrid before it: no table upload
        s1 = tb.transfer_stats()
        out = tb.state_get_grown(idx)
        delta = {k: s1[k] - s0[k] for k in s0}
        return out, delta, nblk
    finally:
        tb.close()


def test_resident_fill_loop_moves_no_state():
    host, dh, nb_h = fill_loop(False)
    print(f"[gpu] host-data fill loop: blocks {nb_h}, traffic {dh}", flush=True)
    dev, dd, nb_d = fill_loop(True)
    print(f"[gpu] resident fill loop: blocks {nb_d}, traffic {dd}", flush=True)
    assert nb_h == nb_d and len(set(nb_h)) > 1         # the grid changes from step to step
    assert np.array_equal(dev.view(np.int64), host.view(np.int64))
    # no state copies between the initial upload and the final download ...
    assert dd["to_device"] == 0 and dd["to_host"] == 0
    # ... where the host-data loop copies the whole state in and out at every fill and new-block fill
    assert dh["to_device"] > 0 and dh["to_host"] > 0
    # tables: at most once per regrid (the closing fill without a regrid uploads nothing)
    steps = len(nb_d)
    assert 1 <= dd["topology_uploads"] <= steps and dd["surroundings_uploads"] == dd["topology_uploads"]
    assert dh["topology_uploads"] == dd["topology_uploads"]
