"""Ghost fill of face-centred data (FillPatch::fillBoundary on a face TreeData,
SayakaFillPatch.cpp:17-89, 620-700; prolongFaceBlockData_*, SayakaInterpPatch.cpp:447-591):
the oracle's restatement (oracle/fill_oracle.py FaceFill) against the reference's own output.

* Golden: tests/golden/fface{2,3}.npz hold the reference's ghost faces after fillBoundary on
  face data of every direction of a periodic 2D quadtree (8^2 blocks, three components with
  injection / centred-linear / limited prolongation, 2 and 1 layers) and a periodic 3D octree
  (4^3 blocks, one component per direction, centred linear / limited / injection, 2 layers).
  The valid faces come from a hash of the global face position, so the two copies of a face
  shared by two blocks agree (the reference's result then depends on no block order).
  Bit-exact, stale ghosts included.
* The product tree reproduces the recorded periodic tree tables (surroundings wrap around).
* Host checks: the oracle refuses a tree with walls, as the reference exits there.
"""
from __future__ import annotations

import numpy as np
import pytest

import golden_cases as GC
from oracle import fill_oracle as F
from sayakaamr_amd import workloads as W
from sayakaamr_amd.tree import AmrTree


@pytest.mark.parametrize("name", ["fface2", "fface3"])
def test_product_periodic_tree_matches_reference_tables(name, golden):
    g = golden(name)
    spec = GC.face_spec(name)
    t = W.build_tree(AmrTree(spec.ndim, max_blocks=4000), spec, GC.face_criterion(name))
    for k in GC.TREE_KEYS:
        if "tree_" + k in g:
            assert np.array_equal(t[k], g["tree_" + k]), k
    # periodic: no physical surrounding anywhere
    assert (t["surr"] > F.PHYS_BNDRY).all()


@pytest.mark.parametrize("name", ["fface2", "fface3"])
def test_face_oracle_matches_reference_fill(name, golden):
    g = golden(name)
    spec = GC.face_spec(name)
    tree = GC.tree_of(g)
    nd, nb = spec.ndim, spec.nb
    for fd in range(nd):
        prolong = GC.face_prolong(name, fd)
        mask = GC.face_ghost_mask(nd, nb, GC.FILL_NGROW, fd)
        for nlayer in GC.FACE_CASES[name][5]:
            data = GC.face_input(name, tree, fd)
            st = [data[c].copy() for c in range(data.shape[0])]
            ff = F.FaceFill(tree, nb, nd, GC.FILL_NGROW, spec.maxlevel, fd)
            ff.fill(st, 0, len(st), nlayer, prolong)
            out = np.stack(st)
            ref = g[f"ghosts_d{fd}_nlayer{nlayer}"]
            assert np.array_equal(out[:, :, mask].view(np.int64), ref.view(np.int64)), (fd, nlayer)
            assert np.array_equal(out[:, :, ~mask], data[:, :, ~mask])


def test_face_oracle_refuses_walls():
    spec = GC.fill_spec("fill2")
    t = W.build_tree(AmrTree(2, max_blocks=4000), spec, GC.fill_criterion("fill2"))
    n = t["level"].shape[0]
    st = [np.zeros((n,) + GC.face_shape(2, spec.nb, GC.FILL_NGROW, 0))]
    ff = F.FaceFill(t, spec.nb, 2, GC.FILL_NGROW, spec.maxlevel, 0)
    with pytest.raises(ValueError, match="wall"):
        ff.fill(st, 0, 1, 1, [0])
