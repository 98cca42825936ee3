"""Chombo-format HDF5 output (WriteDataHdf5, SayakaWriterHdf5.cpp:17-530) written by the
build's own HDF5 writer (sayakaamr_amd/csrc/host/h5lite.cpp, through sy_write_chombo).

The checker is the image's HDF5 library (libhdf5 1.10 under /opt/conda/lib, an independent
implementation of the file format; the tests skip where it is absent), via
oracle/chombo_oracle.py:
* the reference's writer restated call for call over that library writes the same tree and
  state; both files are read back through the library and must agree object by object:
  groups, links, attribute types and bytes, dataset types, shapes and bytes;
* the restatement itself is pinned by what the geometry fixes: every box maps back to its
  block's corners, every level's prob_domain covers the problem box, the data are the state's
  valid cells block by block, and the attributes carry the tree's values;
* a tree with more levels than a default symbol-table node holds (root with 10+ links),
  2D and 3D, component names (given, defaulted, longer than the 64-byte field), and the
  error paths.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from oracle import chombo_oracle as CO
from sayakaamr_amd import output
from sayakaamr_amd import workloads as W
from sayakaamr_amd.tree import AmrTree

pytestmark = pytest.mark.skipif(CO.HDF5 is None, reason="no HDF5 library in this image to read the files back")


def make_case(nd, maxlevel, nb, ngrow=1, ncomp=2, seed=0, crit=None):
    spec = W.small_spec(nd, maxlevel, nb=nb)
    if crit is None:
        crit = (W.ring_criterion(centre=(0.5, 0.5), r0=0.2, half=0.04) if nd == 2
                else W.shell_criterion(centre=(0.4, 0.5, 0.5), r0=0.18))
    tb = AmrTree(nd, max_blocks=8000)
    t = W.build_tree(tb, spec, crit)
    n = t["level"].shape[0]
    g = ngrow
    shp = (nb + 2 * g,) * 3 if nd == 3 else (1, nb + 2 * g, nb + 2 * g)
    state = np.random.default_rng(seed).standard_normal((ncomp, n) + shp)
    return tb, t, spec, state


def write_both(tmp_path, tb, t, spec, state, ngrow, **meta):
    nb, nd = spec.nb, spec.ndim
    prod = str(tmp_path / "product.h5")
    ref = str(tmp_path / "reference.h5")
    output.write_chombo(tb, CO.pack_levels(t, nb, nd, state, ngrow), state.shape[0], prod, **meta)
    CO.write_reference(ref, t, nb, nd, state, ngrow, **meta)
    return CO.HDF5.dump(prod), CO.HDF5.dump(ref)


@pytest.mark.parametrize("nd,nb", [(2, 8), (3, 4)])
def test_product_file_equals_reference_writer(tmp_path, nd, nb):
    tb, t, spec, state = make_case(nd, 4 if nd == 2 else 3, nb, ngrow=2, ncomp=3, seed=nd)
    p, r = write_both(tmp_path, tb, t, spec, state, 2, time=0.75, dt=0.125, step=42, names=["pres", "", "vof"])
    assert sorted(p) == sorted(r)
    for k in r:
        assert p[k] == r[k], k


def _f64(raw):
    return float(np.frombuffer(raw, np.float64)[0])


def _i32(raw):
    return np.frombuffer(raw, np.int32)


@pytest.mark.parametrize("nd,nb", [(2, 8), (3, 4)])
def test_reference_writer_restatement_is_pinned_by_geometry(tmp_path, nd, nb):
    tb, t, spec, state = make_case(nd, 4 if nd == 2 else 3, nb, ngrow=1, ncomp=2, seed=7)
    _, r = write_both(tmp_path, tb, t, spec, state, 1, time=1.5, dt=0.25, step=3)
    root = r["/"][1]
    finest = int(t["level"].max())
    assert _i32(root["num_levels"][2])[0] == finest
    assert _i32(root["num_components"][2])[0] == 2
    assert _i32(root["iteration"][2])[0] == 3 and _f64(root["time"][2]) == 1.5
    assert root["component_0"][2].rstrip(b"\0") == b"comp0" and root["component_1"][2].rstrip(b"\0") == b"comp1"
    problo = np.frombuffer(root["prob_lo"][2], np.float64)
    assert np.array_equal(problo, np.array(spec.lo[:nd]))
    assert _i32(r["/Chombo_global"][1]["SpaceDim"][2])[0] == nd
    domain = np.array(spec.hi[:nd]) - np.array(spec.lo[:nd])
    off = 0
    for il in range(finest):
        lv = r[f"/level_{il}"][1]
        blocks = np.nonzero(t["level"] == il + 1)[0]
        dx = t["length"][blocks[0], 0] / nb
        assert _f64(lv["dx"][2]) == dx and _f64(lv["dt"][2]) == 0.25 and _i32(lv["ref_ratio"][2])[0] == 2
        dom = _i32(lv["prob_domain"][2])
        # cells of the whole domain at this level: hi + 1 per direction
        assert np.array_equal(dom[nd:] + 1, np.rint(domain / dx).astype(np.int64)) and not dom[:nd].any()
        boxes = _i32(r[f"/level_{il}/boxes"][3]).reshape(-1, 2 * nd)
        assert boxes.shape[0] == blocks.size
        assert np.allclose(boxes[:, :nd] * dx + problo, t["lo"][blocks, :nd], atol=1e-12)
        assert np.allclose((boxes[:, nd:] + 1) * dx + problo, t["hi"][blocks, :nd], atol=1e-12)
        data = np.frombuffer(r[f"/level_{il}/data:datatype=0"][3], np.float64)
        assert data.size == blocks.size * 2 * nb ** nd
        g = 1
        for q, b in enumerate(blocks):
            for c in range(2):
                valid = state[c, b][(slice(g, g + nb) if nd == 3 else slice(0, 1)), g:g + nb, g:g + nb]
                assert np.array_equal(data[(q * 2 + c) * nb ** nd:(q * 2 + c + 1) * nb ** nd], valid.ravel())
        off += data.size
        assert r[f"/level_{il}/data_attributes"][1]["comps"][2] == np.int32(2).tobytes()
    assert off == CO.pack_levels(t, nb, nd, state, 1).size


def test_many_levels_and_long_names(tmp_path):
    # refine the blocks holding one point, down to level 11: 12 links under the root
    pt = np.array([0.3, 0.6, 0.0])

    def crit(tree, nb, ndim, blocks):
        lo, hi = tree["lo"][blocks, :2], tree["hi"][blocks, :2]
        return np.all((lo <= pt[:2]) & (pt[:2] < hi), axis=1)
    tb, t, spec, state = make_case(2, 11, 4, ngrow=1, ncomp=2, seed=3, crit=crit)
    assert int(t["level"].max()) == 11
    p, r = write_both(tmp_path, tb, t, spec, state, 1, names=["x" * 80, "short"])
    assert len(r["/"][2]) == 12
    assert sorted(p) == sorted(r)
    for k in r:
        assert p[k] == r[k], k
    assert p["/"][1]["component_0"][2] == b"x" * 64


def test_writer_errors(tmp_path):
    tb, t, spec, state = make_case(2, 3, 8, ngrow=1, ncomp=1)
    packed = CO.pack_levels(t, 8, 2, state, 1)
    with pytest.raises(ValueError, match="packed data"):
        output.write_chombo(tb, packed[:-1], 1, str(tmp_path / "a.h5"))
    with pytest.raises(output.WriterError, match="cannot open"):
        output.write_chombo(tb, packed, 1, str(tmp_path / "no_such_dir" / "a.h5"))
