"""Ghost fill of device-resident face data (sayaka_dev_face_fill, the device
FillPatch::fillBoundary on a face TreeData) against the reference's own output and the oracle.

* tests/golden/fface{2,3}.npz: the reference's ghost faces after fillBoundary on face data of
  every direction of a periodic 2D quadtree (8^2 blocks; injection / centred linear / limited;
  2 and 1 layers) and a periodic 3D octree (4^3 blocks, 2 layers).  The device fill runs on
  the product tree built from the same flags.  Bit-exact, including the ghosts the reference
  leaves untouched; valid faces unchanged.
* 3D 8^3 blocks with 3 ghost layers (an odd nlayer below ngrow, whose centred-linear mid-face
  average reads the next layer): against oracle/fill_oracle.py on the same input.
* Error paths (host-side checks, no launch): a tree with walls, an odd nlayer equal to ngrow
  with centred-linear prolongation, a face direction beyond the tree's dimension.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

import golden_cases as GC
from oracle import fill_oracle as F
from sayakaamr_amd import device as D
from sayakaamr_amd import workloads as W
from sayakaamr_amd.tree import AmrTree

pytestmark = pytest.mark.gpu


def setup(spec, crit, max_blocks=4000):
    t = W.build_tree(AmrTree(spec.ndim, max_blocks=max_blocks), spec, crit)
    dev = D.MacSolverGPU(spec.ndim, spec.nb, max_blocks, device=0)
    dev.prepare(t, spec)
    dev.set_surroundings(t)
    return t, dev


def run_fill(dev, data, ngrow, fd, nlayer, prolong, scomp=0, nc=None):
    n = data.shape[1]
    st = D.DeviceState(dev, data.shape[0], ngrow, n, fdir=fd)
    try:
        assert st.shape == data.shape
        st.upload(data)
        st.fill(nlayer, prolong, scomp=scomp, nc=nc)
        return st.download()
    finally:
        st.close()


@pytest.mark.parametrize("name", ["fface2", "fface3"])
def test_device_face_fill_matches_reference(name, golden):
    g = golden(name)
    spec = GC.face_spec(name)
    t, dev = setup(spec, GC.face_criterion(name))
    n = t["level"].shape[0]
    try:
        for fd in range(spec.ndim):
            prolong = GC.face_prolong(name, fd)
            mask = GC.face_ghost_mask(spec.ndim, spec.nb, GC.FILL_NGROW, fd)
            for nlayer in GC.FACE_CASES[name][5]:
                print(f"[gpu] face fill {name}: {n} blocks, dir {fd}, nlayer {nlayer}", flush=True)
                data = GC.face_input(name, t, fd)
                out = run_fill(dev, data, GC.FILL_NGROW, fd, nlayer, prolong)
                ref = g[f"ghosts_d{fd}_nlayer{nlayer}"]
                assert np.array_equal(out[:, :, mask].view(np.int64), ref.view(np.int64)), (fd, nlayer)
                assert np.array_equal(out[:, :, ~mask], data[:, :, ~mask])
    finally:
        dev.close()


def _periodic_8cube():
    spec = dataclasses.replace(W.small_spec(3, 3, nb=8), bc=(1,) * 6, ngrow=3)
    return spec, W.shell_criterion(centre=(0.3, 0.6, 0.5), r0=0.2)


def _consistent_faces(t, nb, ngrow, fd, ncomp, spec, seed):
    """Face data whose shared faces agree: valid faces from a table over global face positions
    (periodic wrap) of each block's level, ghosts a sentinel."""
    n = t["level"].shape[0]
    shp = GC.face_shape(3, nb, ngrow, fd)
    out = np.full((ncomp, n) + shp, -7.0)
    rng = np.random.default_rng(seed)
    tables = {}
    for b in range(n):
        lvl = int(t["level"][b])
        dx = t["length"][b] / nb
        base = np.rint((t["lo"][b] - np.array(spec.lo)) / dx).astype(np.int64)
        ncells = np.array(spec.ndiv, dtype=np.int64) * (1 << (lvl - 1)) * nb
        if lvl not in tables:
            tables[lvl] = rng.uniform(-1, 1, (ncomp,) + tuple(int(x) for x in ncells[::-1]))
        ext = [nb + (d == fd) for d in range(3)]
        idx = [(base[d] + np.arange(ext[d])) % ncells[d] for d in range(3)]
        vals = tables[lvl][:, idx[2][:, None, None], idx[1][None, :, None], idx[0][None, None, :]]
        out[:, b, ngrow:ngrow + ext[2], ngrow:ngrow + ext[1], ngrow:ngrow + ext[0]] = vals
    return out


@pytest.mark.parametrize("fd", [0, 1, 2])
def test_device_face_fill_3d_8cube_matches_oracle(fd):
    spec, crit = _periodic_8cube()
    t, dev = setup(spec, crit)
    n = t["level"].shape[0]
    prolong = [1, 2, 0]
    data = _consistent_faces(t, 8, 3, fd, 3, spec, seed=20 + fd)
    print(f"[gpu] face fill 3D 8^3: {n} blocks, dir {fd}", flush=True)
    try:
        out3 = run_fill(dev, data, 3, fd, 3, [0, 0, 0])          # odd nlayer == ngrow: injection only
        out1 = run_fill(dev, data, 3, fd, 1, prolong)            # odd nlayer below ngrow
        out2 = run_fill(dev, data, 3, fd, 2, prolong[1:], scomp=1, nc=2)
    finally:
        dev.close()
    ff = F.FaceFill(t, 8, 3, 3, spec.maxlevel, fd)
    for out, nl, pr, sc, nc in ((out3, 3, [0, 0, 0], 0, 3), (out1, 1, prolong, 0, 3), (out2, 2, prolong, 1, 2)):
        ref = [data[c].copy() for c in range(3)]
        ff.fill(ref, sc, nc, nl, pr)
        assert np.array_equal(out.view(np.int64), np.stack(ref).view(np.int64)), (nl, sc)


def test_device_face_fill_rejects_walls_and_bad_arguments():
    spec = GC.fill_spec("fill2")
    t, dev = setup(spec, GC.fill_criterion("fill2"))
    n = t["level"].shape[0]
    try:
        data = np.zeros((1, n) + GC.face_shape(2, spec.nb, GC.FILL_NGROW, 0))
        with pytest.raises(D.DeviceError, match="physical boundaries"):
            run_fill(dev, data, GC.FILL_NGROW, 0, 1, [0])
    finally:
        dev.close()
    spec = GC.face_spec("fface2")
    t, dev = setup(spec, GC.face_criterion("fface2"))
    n = t["level"].shape[0]
    try:
        data = np.zeros((1, n) + GC.face_shape(2, spec.nb, 1, 1))
        with pytest.raises(D.DeviceError, match="odd nlayer"):
            run_fill(dev, np.zeros((1, n) + GC.face_shape(2, spec.nb, 3, 1)), 3, 1, 3, [1])
        with pytest.raises(D.DeviceError, match="ngrow >= 2"):
            run_fill(dev, data, 1, 1, 1, [2])
        st = D.DeviceState(dev, 1, 2, n, fdir=0)
        try:
            with pytest.raises(D.DeviceError, match="face direction"):
                dev.face_fill(st.ptr, 1, n, 2, 2, 1, [0])
        finally:
            st.close()
    finally:
        dev.close()
