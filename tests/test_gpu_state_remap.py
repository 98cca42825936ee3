"""Regrid remap of a registered state on the device (sayaka_dev_state_remap: AmrTree::regrid's
packing and initNewBlockDataAfterRegrid, SayakaRegrid.cpp:530-591, 987-1105) against the
reference's own output.

tests/golden/{s,f}remap{2,3}.npz: regrids of a moving ring (2D, 8^2 blocks) / shell (3D, 4^3
blocks) with a registered cell state (3 / 2 components: injection, centred linear, limited;
every wall its own code and BC mix) and face states of every direction on a periodic tree.
The product tree replays the regrids with the recorded flags; the device state on the old tree
is remapped into a state on the new tree.  Bit-exact on the blocks the remap writes (new
blocks, parents of new blocks: the fixture), and every other block holds its moved data.
Also: 3D 8^3 blocks against the oracle, and the host-side error paths.
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


def run_remap(dev, spec, before, after, pre, origin, fd, prolong, bcmap):
    ncomp = pre.shape[0]
    old = D.DeviceState(dev, ncomp, GC.FILL_NGROW, before["level"].shape[0], fdir=fd)
    new = D.DeviceState(dev, ncomp, GC.FILL_NGROW, after["level"].shape[0] + 3, fdir=fd)
    try:
        old.upload(pre)
        dev.prepare(after, spec)
        dev.set_surroundings(after)
        new.remap_from(old, origin, prolong, bcmap)
        return new.download()[:, :after["level"].shape[0]]
    finally:
        old.close()
        new.close()


@pytest.mark.parametrize("name", ["sremap2", "sremap3", "fremap2", "fremap3"])
def test_device_remap_matches_reference(name, golden):
    g = golden(name)
    spec = GC.remap_spec(name)
    dirs = GC.REMAP_CASES[name][3]
    dev = D.MacSolverGPU(spec.ndim, spec.nb, 4000, device=0)
    try:
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
                print(f"[gpu] state remap {name} {tag} step {step}: {before['level'].shape[0]} -> "
                      f"{after['level'].shape[0]} blocks, {int((origin < 0).sum())} new", flush=True)
                post = run_remap(dev, spec, before, after, pre, origin, fd, GC.remap_prolong(name, fd),
                                 GC.remap_bcmaps(name) if fd is None else None)
                touched = GC.remap_touched(after, origin)
                assert np.array_equal(post[:, touched].view(np.int64), g[f"{tag}_step{step}_post"].view(np.int64)), \
                    (fd, step)
                keep = ~touched
                assert np.array_equal(post[:, keep], pre[:, origin[keep]])
    finally:
        dev.close()


def test_device_remap_3d_8cube_matches_oracle():
    spec = dataclasses.replace(W.small_spec(3, 3, nb=8), bc=GC.FILL_WALLS, ngrow=2)
    tb = AmrTree(3, max_blocks=4000)
    W.build_tree(tb, spec, W.shell_criterion(centre=(0.4, 0.5, 0.5), r0=0.18))
    before = tb.export()
    refine, coarsen = GC.remap_flags(before, spec, W.shell_criterion(centre=(0.5, 0.5, 0.5), r0=0.18))
    tb.set_refine_flags(refine)
    tb.set_coarsen_flags(coarsen)
    tb.regrid()
    after = tb.export()
    origin = tb.regrid_origin()
    assert (origin < 0).any()
    pre = np.random.default_rng(3).standard_normal((3, before["level"].shape[0], 12, 12, 12))
    prolong = [0, 1, 2]
    bcmaps = [{code: ((q + c) % 3, 0.5 * q - c) for q, code in enumerate(GC.FILL_WALLS)} for c in range(3)]
    print(f"[gpu] state remap 3D 8^3: {before['level'].shape[0]} -> {after['level'].shape[0]} blocks", flush=True)
    dev = D.MacSolverGPU(3, 8, 4000, device=0)
    try:
        post = run_remap(dev, spec, before, after, pre, origin, None, prolong, bcmaps)
    finally:
        dev.close()
    ref = F.remap_state(after, origin, pre, 8, 3, 2, spec.maxlevel, prolong, bcmaps)
    assert np.array_equal(post.view(np.int64), ref.view(np.int64))


def test_device_remap_errors():
    spec = GC.remap_spec("sremap2")
    tb = AmrTree(2, max_blocks=4000)
    t = W.build_tree(tb, spec, GC.remap_criterion("sremap2", 0))
    n = t["level"].shape[0]
    dev = D.MacSolverGPU(2, spec.nb, 4000, device=0)
    try:
        dev.prepare(t, spec)
        dev.set_surroundings(t)
        a = D.DeviceState(dev, 1, 2, n)
        b = D.DeviceState(dev, 1, 2, n)
        f0 = D.DeviceState(dev, 1, 2, n, fdir=0)
        f1 = D.DeviceState(dev, 1, 2, n, fdir=0)
        try:
            ident = np.arange(n, dtype=np.int32)
            partial = {code: (0, 0.0) for code in GC.FILL_WALLS[:2]}
            with pytest.raises(D.DeviceError, match="no BC mapping"):
                b.remap_from(a, ident, [1], [partial])
            with pytest.raises(D.DeviceError, match="periodic"):
                f1.remap_from(f0, ident, [1])
            bad = ident.copy()
            bad[3] = n + 7
            with pytest.raises(D.DeviceError, match="origin"):
                b.remap_from(a, bad, [0], [{c: (0, 0.0) for c in GC.FILL_WALLS}])
            with pytest.raises(ValueError, match="differ"):
                f1.remap_from(a, ident, [0])
        finally:
            for x in (a, b, f0, f1):
                x.close()
    finally:
        dev.close()


@pytest.mark.parametrize("name", ["sremap2", "sremap3", "fremap2"])
def test_host_registered_state_matches_reference(name, golden):
    """The same regrids through the host C++ surface a driver uses: a state registered with
    the AmrTree (TreeStateData with its FillPatch / BCPatch), AmrTree::regrid moving its blocks
    on the host and AmrTree::initNewBlockDataAfterRegrid filling the new ones on the GPU."""
    g = golden(name)
    spec = GC.remap_spec(name)
    dirs = GC.REMAP_CASES[name][3]
    for fd in (dirs if dirs is not None else (None,)):
        tag = "c" if fd is None else f"d{fd}"
        tb = AmrTree(spec.ndim, max_blocks=4000)
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
        for step in range(GC.REMAP_CASES[name][5]):
            before = tb.export()
            pre = GC.remap_input(name, before, step, fd)
            tb.state_set_grown(idx, pre)
            tb.set_refine_flags(g[f"{tag}_step{step}_refine"])
            tb.set_coarsen_flags(g[f"{tag}_step{step}_coarsen"])
            tb.regrid()
            print(f"[gpu] host registered state {name} {tag} step {step}", flush=True)
            tb.init_new_data()
            after = tb.export()
            origin = tb.regrid_origin()
            post = tb.state_get_grown(idx)
            touched = GC.remap_touched(after, origin)
            assert np.array_equal(post[:, touched].view(np.int64), g[f"{tag}_step{step}_post"].view(np.int64)), \
                (fd, step)
            keep = ~touched
            assert np.array_equal(post[:, keep], pre[:, origin[keep]])
