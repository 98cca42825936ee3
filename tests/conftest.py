"""Shared test set-up: import paths, the `gpu` marker and the golden fixtures.

CPU tests (`-m "not gpu"`) cover the host tree against the reference's recorded tables, the C
oracle against the golden vectors, invariants and closed forms, and the C ABI's exports.
GPU tests (`-m gpu`) run the HIP solver on the device through the C ABI and compare it with
the oracle and the golden vectors.
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.join(ROOT, "tests")
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(TESTS, "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the solver on the GPU")


def _ensure_built():
    """Build the native libraries only when they are missing (build() rebuilds on demand)."""
    from oracle import oraclebind
    from sayakaamr_amd import _native as N

    if not os.path.exists(oraclebind.LIB):
        oraclebind.build()
    if not os.path.exists(N.HOST_LIB) or not os.path.exists(N.DEV_LIB):
        from sayakaamr_amd import build as B

        if not os.path.exists(N.HOST_LIB):
            B.build_host()
        if not os.path.exists(N.DEV_LIB):
            B.build_dev()


_ensure_built()


@pytest.fixture(scope="session")
def golden():
    cache = {}

    def load(name):
        if name not in cache:
            with np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False) as z:
                cache[name] = {k: z[k] for k in z.files}
        return cache[name]

    return load
