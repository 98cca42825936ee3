"""A C++ driver written against the MacSolver / AmrTree / TreeData headers (the reference's
call surface, Test7.h:411-418, 612-653), run as its own process on the GPU.

This file sorts first among the GPU tests on purpose: the bottom solve is a cooperative
launch, and the HIP runtime refuses a cooperative launch in a second process while another
process on the card holds one (the pytest process would, once any solve had run in it)."""
from __future__ import annotations

import os
import subprocess

import pytest

from sayakaamr_amd import _native as N

pytestmark = pytest.mark.gpu


def test_cpp_driver():
    exe = os.path.join(N.LIBDIR, "sayaka_c1_solve")
    print("[gpu] C++ driver", exe, flush=True)
    r = subprocess.run([exe, "4", "0"], capture_output=True, text=True, timeout=100)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stderr
    kv = dict(tok.split("=") for line in r.stdout.splitlines() for tok in line.split() if "=" in tok)
    assert kv["blocks"] == "4864" and kv["leaves"] == "4264"
    assert kv["fp64_status"] == "0" and kv["mixed_status"] == "0"
    assert int(kv["fp64_iters"]) < 50
    assert float(kv["rel_l2_mixed_vs_fp64"]) < 1e-3
