"""General ghost fill of a device-resident cell state (sayaka_dev_state_fill, the device
FillPatch::fillBoundary) against the reference's own output and the oracle.

* tests/golden/fill{2,3}.npz: the reference's ghost cells after fillBoundary (2D 8^2 blocks,
  three components with injection / centred-linear / limited prolongation, 1 and 2 layers; 3D
  4^3 blocks, centred linear and limited), every wall its own physical code and BC mix.  The
  device fill runs on the product tree built from the same flags.  Bit-exact, including the
  ghosts the reference leaves untouched.
* 3D 8^3 blocks (the 3D solve's block size): against oracle/fill_oracle.py on the same input.
* Error paths (host-side checks, no launch): a wall code without a BC mapping, nlayer > ngrow.
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
    tb = AmrTree(spec.ndim, max_blocks=max_blocks)
    t = W.build_tree(tb, spec, crit)
    dev = D.MacSolverGPU(spec.ndim, spec.nb, max_blocks, device=0)
    dev.prepare(t, spec)
    dev.set_surroundings(t)
    return t, dev


@pytest.mark.parametrize("name", ["fill2", "fill3"])
def test_device_fill_matches_reference(name, golden):
    g = golden(name)
    spec = GC.fill_spec(name)
    t, dev = setup(spec, GC.fill_criterion(name))
    n = t["level"].shape[0]
    comps = GC.FILL_COMPS[name]
    mask = GC.ghost_mask(spec.ndim, spec.nb, GC.FILL_NGROW)
    try:
        for nlayer in GC.FILL_CASES[name][5]:
            print(f"[gpu] state fill {name}: {n} blocks, nlayer {nlayer}", flush=True)
            data = GC.fill_input(name, n)
            st = D.DeviceState(dev, data.shape[0], GC.FILL_NGROW, n)
            try:
                st.upload(data)
                st.fill(nlayer, comps["prolong"], GC.fill_bcmaps(name))
                out = st.download()
            finally:
                st.close()
            ref = g[f"ghosts_nlayer{nlayer}"]
            assert np.array_equal(out[:, :, mask].view(np.int64), ref.view(np.int64)), nlayer
            assert np.array_equal(out[:, :, ~mask], data[:, :, ~mask])
    finally:
        dev.close()


def test_device_fill_3d_8cube_matches_oracle():
    spec = dataclasses.replace(W.small_spec(3, 3, nb=8), bc=GC.FILL_WALLS, ngrow=2)
    crit = W.shell_criterion(centre=(0.4, 0.5, 0.5), r0=0.18)
    t, dev = setup(spec, crit)
    n = t["level"].shape[0]
    print(f"[gpu] state fill 3D 8^3: {n} blocks", flush=True)
    rng = np.random.default_rng(11)
    data = rng.standard_normal((3, n, 12, 12, 12))
    prolong = [0, 1, 2]
    bcmaps = [{code: ((q + c) % 3, 0.5 * q - c) for q, code in enumerate(GC.FILL_WALLS)} for c in range(3)]
    try:
        st = D.DeviceState(dev, 3, 2, n)
        try:
            st.upload(data)
            st.fill(2, prolong, bcmaps)
            out = st.download()
            # a second fill of components 1..2 only, one layer: component 0 stays as it was
            st.fill(1, prolong[1:], bcmaps[1:], scomp=1, nc=2)
            out2 = st.download()
        finally:
            st.close()
    finally:
        dev.close()
    ref = [data[c].copy() for c in range(3)]
    sf = F.StateFill(t, 8, 3, 2, spec.maxlevel)
    sf.fill(ref, 0, 3, 2, prolong, bcmaps)
    assert np.array_equal(out.view(np.int64), np.stack(ref).view(np.int64))
    sf.fill(ref, 1, 2, 1, prolong, bcmaps)
    assert np.array_equal(out2.view(np.int64), np.stack(ref).view(np.int64))


def test_device_fill_rejects_unmapped_wall_and_bad_layers():
    spec = GC.fill_spec("fill2")
    t, dev = setup(spec, GC.fill_criterion("fill2"))
    n = t["level"].shape[0]
    try:
        st = D.DeviceState(dev, 1, 2, n)
        try:
            partial = {code: (F.NEUMANN, 0.0) for code in GC.FILL_WALLS[:3]}
            with pytest.raises(D.DeviceError, match="no BC mapping"):
                st.fill(2, [1], [partial])
            full = {code: (F.NEUMANN, 0.0) for code in GC.FILL_WALLS}
            with pytest.raises(D.DeviceError, match="nlayer"):
                st.fill(3, [1], [full])
        finally:
            st.close()
    finally:
        dev.close()
