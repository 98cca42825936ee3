"""Synthetic inputs of the benchmark workloads: the properties they are defined by."""
from __future__ import annotations

import subprocess
import sys
import os

import numpy as np

import parity_util as U
from sayakaamr_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_c1_inputs_properties(golden):
    spec, t, tree, inp = U.c1_case()
    n = tree["level"].shape[0]
    leaf = tree["node_type"] == 1
    vf, rhs = inp["vf"], inp["rhs"]
    assert vf.shape == (n, 512) and np.all((vf >= 0) & (vf <= 1))
    # cells inside the sphere are solid, far cells are fluid
    x, y, z = W.cell_centres(tree, 8, 3)
    r = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2).reshape(n, -1)
    assert np.all(vf[r < 0.15] == 0) and np.all(vf[r > 0.25] == 1)
    # rhs vanishes on solid cells and has zero volume-weighted mean over the fluid
    assert np.all(rhs[vf == 0] == 0)
    dv = (tree["length"][:, 0] / 8) ** 3
    mean = np.sum((dv[:, None] * vf * rhs)[leaf])
    assert abs(mean) <= 1e-12 * np.sum(np.abs(dv[:, None] * vf * rhs)[leaf])
    # density ratio 1000: B in [1e-3, 1]
    for d in range(3):
        b = inp["b"][d][leaf]
        assert b.min() >= 1e-3 - 1e-15 and b.max() <= 1.0
    g = golden("c1_summary")
    assert rhs.sum() == g["rhs_checksum"][0]


def test_face_velocity_independent_of_numbering():
    """The seeded face velocity is keyed on (level, global face, dir), not on block ids."""
    spec, t, tree, inp = U.c1_case()
    u1 = W.face_velocity(tree, spec, 7)
    u2 = W.face_velocity(tree, spec, 7)
    u3 = W.face_velocity(tree, spec, 8)
    assert all(np.array_equal(a, b) for a, b in zip(u1, u2))
    assert not np.array_equal(u1[0], u3[0])
    assert all(np.all(np.abs(a) <= 1.0) for a in u1)


def test_pin_cell_contains_point():
    spec, t, tree, inp = U.c1_case()
    b, c = W.pin_cell(tree, spec)
    assert tree["node_type"][b] == 1
    dx = tree["length"][b, 0] / 8
    i, j, k = c % 8, (c // 8) % 8, c // 64
    cc = tree["lo"][b] + (np.array([i, j, k]) + 0.5) * dx
    assert np.all(np.abs(cc - np.array(W.PIN_POINT)) <= 0.5 * dx + 1e-15)


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "8"], capture_output=True, text=True)
    assert r.returncode != 0 and "one GPU" in r.stderr
