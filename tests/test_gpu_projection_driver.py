"""The projection driver built against this build's include/sayaka, on the GPU, against the same
driver built against the reference's sources: every block of the pressure (after the solve, the
fill of its ghosts and mg_flux_level's fill), of the corrected velocity and of the flux
(mg_flux_level into the caller's face data, restricted fine fluxes on the parents' faces
included) bit for bit (a 64-bit hash per block and array, tests/golden/proj_step.npz)."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import golden_cases as GC
from sayakaamr_amd import _native as N

pytestmark = pytest.mark.gpu


def test_projection_driver_matches_reference_build(tmp_path, golden):
    g = golden("proj_step")
    exe = os.path.join(N.LIBDIR, "sayaka_projection_step")
    out = str(tmp_path / "gpu.bin")
    print("[gpu] projection driver", exe, flush=True)
    r = subprocess.run([exe, out, str(GC.PROJ_MAXLEVEL)], capture_output=True, text=True, timeout=100)
    print(r.stdout[-500:], flush=True)
    assert r.returncode == 0, r.stderr
    info, arrs = GC.read_projection_dump(out)
    assert info["nblk"] == int(g["nblk"]) and info["finest"] == int(g["finest"])
    assert np.array_equal(info["tower"], g["tower"])
    bad = {}
    for name, a in arrs.items():
        h = GC.block_hashes(a)
        miss = np.nonzero(h != g[f"h_{name}"])[0]
        if miss.size:
            bad[name] = (int(miss.size), miss[:8].tolist(), float(np.linalg.norm(a)), float(g[f"l2_{name}"][0]))
    assert not bad, bad
