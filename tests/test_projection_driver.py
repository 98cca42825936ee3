"""The projection driver (sayakaamr_amd/csrc/drivers/projection_step.cpp) -- one pressure
projection written in the reference flow demo's idioms and call order (Test7.h:285-480,
580-653) -- builds against both call surfaces: the reference's own sources (oracle/build_ref.sh,
oracle/_ref/ref_projection_step) and this build's include/sayaka (sayakaamr_amd/lib/
sayaka_projection_step).  Here (no GPU): the product build exists, and, where the reference build
is present, a fresh run of it reproduces the golden hashes (tests/golden/proj_step.npz)."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import golden_cases as GC
from sayakaamr_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_EXE = os.path.join(ROOT, "oracle", "_ref", "ref_projection_step")


def test_product_driver_built():
    assert os.access(os.path.join(N.LIBDIR, "sayaka_projection_step"), os.X_OK)


def test_product_driver_compiles_for_2d():
    """The driver's loops are written for NDIM; the headers take SAYAKA_SPACEDIM=2 as well."""
    src = os.path.join(ROOT, "sayakaamr_amd", "csrc", "drivers", "projection_step.cpp")
    inc = os.path.join(ROOT, "include", "sayaka")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-DSAYAKA_SPACEDIM=2", f"-I{inc}", src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(not os.access(REF_EXE, os.X_OK), reason="reference build absent (oracle/_ref)")
def test_reference_build_reproduces_golden(tmp_path, golden):
    g = golden("proj_step")
    out = str(tmp_path / "ref.bin")
    r = subprocess.run([REF_EXE, out, str(GC.PROJ_MAXLEVEL)], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    info, arrs = GC.read_projection_dump(out)
    assert info["nblk"] == int(g["nblk"]) and np.array_equal(info["tower"], g["tower"])
    for name, a in arrs.items():
        assert np.array_equal(GC.block_hashes(a), g[f"h_{name}"]), name
