"""Chombo output of a device-resident cell state: sayaka_dev_pack_levels (valid cells in the
file's order, packed on the device) and the file written from it (DeviceState.write_chombo).

* The packed array equals the oracle's packing (oracle/chombo_oracle.py pack_levels) bit for
  bit, 2D and 3D, several components, ghost widths 1 and 2, state allocated for more blocks
  than the tree holds.
* The file written from the device state equals the reference's writer restated over the
  image's HDF5 library (object by object, read back through that library); skipped where the
  library is absent.
* After a ghost fill the packed valid cells are unchanged (the fill writes ghosts only).
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import chombo_oracle as CO
from sayakaamr_amd import device as D
from sayakaamr_amd import workloads as W
from sayakaamr_amd.tree import AmrTree

pytestmark = pytest.mark.gpu


def setup(nd, nb, maxlevel):
    spec = W.small_spec(nd, maxlevel, nb=nb)
    crit = (W.ring_criterion(centre=(0.5, 0.5), r0=0.2, half=0.04) if nd == 2
            else W.shell_criterion(centre=(0.4, 0.5, 0.5), r0=0.18))
    tb = AmrTree(nd, max_blocks=8000)
    t = W.build_tree(tb, spec, crit)
    dev = D.MacSolverGPU(nd, nb, 8000, device=0)
    dev.prepare(t, spec)
    return tb, t, spec, dev


@pytest.mark.parametrize("nd,nb,maxlevel,ngrow,ncomp", [(2, 8, 4, 1, 3), (3, 8, 3, 2, 2), (3, 4, 3, 1, 1)])
def test_device_pack_matches_oracle(nd, nb, maxlevel, ngrow, ncomp):
    tb, t, spec, dev = setup(nd, nb, maxlevel)
    n = t["level"].shape[0]
    print(f"[gpu] pack levels {nd}D nb {nb}: {n} blocks, ngrow {ngrow}, ncomp {ncomp}", flush=True)
    try:
        st = D.DeviceState(dev, ncomp, ngrow, n + 5)
        try:
            data = np.random.default_rng(nd * 10 + ngrow).standard_normal(st.shape)
            st.upload(data)
            packed = st.pack_levels()
        finally:
            st.close()
    finally:
        dev.close()
    ref = CO.pack_levels(t, nb, nd, data[:, :n], ngrow)
    assert packed.shape == ref.shape
    assert np.array_equal(packed.view(np.int64), ref.view(np.int64))


@pytest.mark.skipif(CO.HDF5 is None, reason="no HDF5 library in this image to read the files back")
@pytest.mark.parametrize("nd,nb", [(2, 8), (3, 8)])
def test_device_state_file_equals_reference_writer(tmp_path, nd, nb):
    tb, t, spec, dev = setup(nd, nb, 4 if nd == 2 else 3)
    n = t["level"].shape[0]
    print(f"[gpu] chombo file {nd}D: {n} blocks", flush=True)
    try:
        dev.set_surroundings(t)
        st = D.DeviceState(dev, 2, 2, n)
        try:
            data = np.random.default_rng(5 + nd).standard_normal(st.shape)
            st.upload(data)
            before = st.pack_levels()
            walls = sorted({int(v) for v in t["surr"].ravel() if v <= -20})
            st.fill(2, [1, 2], [{w: (0, 0.0) for w in walls}] * 2)
            after = st.pack_levels()
            st.write_chombo(tb, str(tmp_path / "dev.h5"), time=2.5, dt=0.5, step=9, names=["p", "q"])
        finally:
            st.close()
    finally:
        dev.close()
    assert np.array_equal(before.view(np.int64), after.view(np.int64))
    CO.write_reference(str(tmp_path / "ref.h5"), t, nb, nd, data, 2, time=2.5, dt=0.5, step=9, names=["p", "q"])
    p, r = CO.HDF5.dump(str(tmp_path / "dev.h5")), CO.HDF5.dump(str(tmp_path / "ref.h5"))
    assert sorted(p) == sorted(r)
    for k in r:
        assert p[k] == r[k], k


def test_face_state_is_refused():
    tb, t, spec, dev = setup(2, 8, 3)
    try:
        st = D.DeviceState(dev, 1, 1, t["level"].shape[0], fdir=0)
        try:
            with pytest.raises(ValueError, match="cell data only"):
                st.pack_levels()
        finally:
            st.close()
    finally:
        dev.close()
