"""sayaka_dev_dot_noalloc -- the allocation-free variant of sayaka_dev_dot (its partials in a
file-scope device array) -- against sayaka_dev_dot through the C ABI: bit-identical results on
two calls in a row, the second over a smaller tower with other data; arguments it must reject."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import parity_util as U
from sayakaamr_amd import device as D

pytestmark = pytest.mark.gpu


def _dev_buf(dev, host):
    p = C.c_void_p()
    a = np.ascontiguousarray(host, dtype=np.float64)
    assert dev.lib.sayaka_dev_alloc(dev.h, a.nbytes, C.byref(p)) == 0
    assert dev.lib.sayaka_dev_memcpy(dev.h, p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0
    return p


def _read2(dev, p):
    out = np.zeros(2)
    assert dev.lib.sayaka_dev_memcpy(dev.h, out.ctypes.data_as(C.c_void_p), p, out.nbytes, 0) == 0
    return out


def test_dot_noalloc_matches_dot():
    spec, _, tree, inp = U.small3_case()
    n = tree["level"].shape[0]
    L = int(tree["level"].max())
    print(f"[gpu] dot_noalloc vs dot: {n} blocks, towers {L} and 2", flush=True)
    dev = D.MacSolverGPU(3, spec.nb, n + 16)
    bufs = []
    try:
        dev.prepare(tree, spec)
        rng = np.random.default_rng(5)
        out = _dev_buf(dev, np.full(2, -7.0))
        bufs.append(out)
        for m, seed in ((L, 1), (2, 2)):          # second call: smaller tower, other data
            x = rng.standard_normal((n, spec.nb ** 3)) * seed
            y = rng.standard_normal((n, spec.nb ** 3))
            dev.upload(D.P, x)
            dev.upload(D.V, y)
            ref = dev.dot(m, D.P, D.V)
            a, b = _dev_buf(dev, x), _dev_buf(dev, y)
            bufs += [a, b]
            assert dev.lib.sayaka_dev_dot_noalloc(dev.h, m, a, b, out) == 0
            got = _read2(dev, out)
            assert got[0] == ref[0] and got[1] == ref[1], (m, got, ref)
            tb = U.tower_blocks(tree, m)
            assert np.isclose(got[0], float(np.sum(x[tb] * y[tb])), rtol=1e-12)
            assert got[1] == np.abs(x[tb]).max()
        assert dev.lib.sayaka_dev_dot_noalloc(dev.h, 0, a, b, out) == -1          # no tower 0
        assert dev.lib.sayaka_dev_dot_noalloc(dev.h, L, a, None, out) == -1       # null operand
    finally:
        for p in bufs:
            dev.lib.sayaka_dev_free(dev.h, p)
        dev.close()
