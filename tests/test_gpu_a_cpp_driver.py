"""A C++ driver written against the MacSolver / AmrTree / TreeData headers (the reference's
call surface, Test7.h:411-418, 612-653), run as its own process on the GPU.

The driver runs as a second process on the GPU beside the pytest process, which has solved
before it when the file order puts it later: the coarsest-level solve is an ordinary launch
with its own grid barrier, so two processes can solve on one GPU.

The driver also writes its FP64 solution through WriteDataHdf5(TreeData) with the raw block
data and geometry beside it; the file must equal, object by object, the reference's writer
restated over the image's HDF5 library on the same tree and data (skipped without that
library).  It then fills the solution's 2 ghost layers with the host FillPatch (centred linear,
Neumann walls); the filled state must equal oracle/fill_oracle.py's fill of the same data on
the same tree, bit for bit."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from oracle import chombo_oracle as CO
from oracle import fill_oracle as F
from sayakaamr_amd import workloads as W
from sayakaamr_amd.tree import AmrTree
from sayakaamr_amd import _native as N

pytestmark = pytest.mark.gpu


def test_cpp_driver(tmp_path):
    exe = os.path.join(N.LIBDIR, "sayaka_c1_solve")
    print("[gpu] C++ driver", exe, flush=True)
    h5 = str(tmp_path / "c1.h5")
    r = subprocess.run([exe, "4", "0", h5], capture_output=True, text=True, timeout=100)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stderr
    kv = dict(tok.split("=") for line in r.stdout.splitlines() for tok in line.split() if "=" in tok)
    assert kv["blocks"] == "4864" and kv["leaves"] == "4264"
    assert kv["fp64_status"] == "0" and kv["mixed_status"] == "0"
    assert int(kv["fp64_iters"]) < 50
    assert float(kv["rel_l2_mixed_vs_fp64"]) < 1e-6   # fluid cells, each solution less its mean
    check_fill(h5)
    check_output(h5)


def check_fill(h5):
    blk = np.fromfile(h5 + ".blocks").reshape(-1, 10)
    n = blk.shape[0]
    # the driver's tree is the benchmark's: same refinement rule, same numbering
    t = W.build_tree(AmrTree(3, max_blocks=8000), W.C1_SPEC, W.shell_criterion())
    assert t["level"].shape[0] == n
    assert np.array_equal(t["level"], blk[:, 0].astype(np.int32))
    assert np.array_equal(t["lo"], blk[:, 1:4]) and np.array_equal(t["hi"], blk[:, 4:7])
    phi = np.fromfile(h5 + ".phi").reshape(n, 8, 8, 8)
    got = np.fromfile(h5 + ".ghost").reshape(n, 12, 12, 12)
    st = np.zeros((n, 12, 12, 12))
    st[:, 2:10, 2:10, 2:10] = phi
    ref = [st]
    F.StateFill(t, 8, 3, 2, W.C1_SPEC.maxlevel).fill(ref, 0, 1, 2, [1], [{W.WALL: (F.NEUMANN, 0.0)}])
    assert np.array_equal(got.view(np.int64), ref[0].view(np.int64))


def check_output(h5):
    blk = np.fromfile(h5 + ".blocks").reshape(-1, 10)
    n = blk.shape[0]
    tree = dict(level=blk[:, 0].astype(np.int32), lo=blk[:, 1:4], hi=blk[:, 4:7], length=blk[:, 7:10])
    phi = np.fromfile(h5 + ".phi").reshape(1, n, 8, 8, 8)
    assert np.isfinite(phi).all()
    if CO.HDF5 is None:
        pytest.skip("no HDF5 library in this image to read the file back")
    ref = h5 + ".ref.h5"
    CO.write_reference(ref, tree, 8, 3, phi, 0, names=["phi"])
    p, r = CO.HDF5.dump(h5), CO.HDF5.dump(ref)
    assert sorted(p) == sorted(r)
    for k in r:
        assert p[k] == r[k], k
