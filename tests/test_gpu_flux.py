"""Face flux on the GPU (k_flux, through sayaka_dev_flux) against the reference's golden
vectors and the oracle: bit-exact on every block of the tower (MacSolver::mg_flux_level,
SayakaMacOperator.cpp:161-239, 929-1025).

The reference also leaves the restricted fine fluxes in the faces of non-leaf parents outside
the tower; the device does not write those blocks, so the comparison covers the tower."""
from __future__ import annotations

import numpy as np
import pytest

from golden_cases import CASES, inputs_of, tower_blocks, tree_of
from oracle import oraclebind as O
from sayakaamr_amd import device as D
from sayakaamr_amd import workloads as W

pytestmark = pytest.mark.gpu

FLUX_FILES = {"c0_ops": "c0_flux", "s3_ops": "s3_flux"}


@pytest.fixture(scope="module", params=sorted(CASES))
def fcase(request, golden):
    spec, bc = CASES[request.param]
    g = golden(request.param)
    gf = golden(FLUX_FILES[request.param])
    tree = tree_of(g)
    inp = inputs_of(g, spec.ndim)
    n = tree["level"].shape[0]
    print(f"[gpu] flux case {request.param}: {n} blocks", flush=True)
    dev = D.MacSolverGPU(spec.ndim, spec.nb, n + 8, device=0)
    dev.prepare(tree, spec)
    dev.set_bc([W.WALL], [bc])
    dev.set_coefs(inp["vf"], inp["af"], inp["b"])
    yield spec, tree, gf, dev
    dev.close()


def test_flux_every_tower_bit_exact(fcase):
    spec, tree, gf, dev = fcase
    dev.upload(D.P, gf["op_x"])
    for m in range(1, spec.maxlevel + 1):
        tb = tower_blocks(tree, m)
        for d in range(spec.ndim):
            f = dev.flux(m, d, D.P)
            assert np.array_equal(f[tb], gf[f"flux_m{m}_d{d}"][tb]), (m, d)
            outside = np.setdiff1d(np.arange(tree["level"].shape[0]), tb)
            assert not np.any(f[outside]), (m, d)


def test_flux_of_solution_bit_exact(fcase):
    spec, tree, gf, dev = fcase
    dev.upload(D.X, gf["sol"])
    tb = tower_blocks(tree, spec.maxlevel)
    for d in range(spec.ndim):
        assert np.array_equal(dev.flux(spec.maxlevel, d, D.X)[tb], gf[f"flux_sol_d{d}"][tb]), d


def test_c1_flux_of_fp64_solution_matches_oracle(golden):
    g = golden("c1_summary")
    spec = W.C1_SPEC
    tree = tree_of(g)
    inp = W.c1_inputs(tree, spec, seed=0)
    n = tree["level"].shape[0]
    dev = D.MacSolverGPU(3, 8, n + 64, device=0)
    try:
        dev.prepare(tree, spec)
        dev.set_bc([W.WALL], [W.NEUMANN])
        dev.set_coefs(inp["vf"], inp["af"], inp["b"])
        dev.upload(D.RHS, inp["rhs"])
        dev.zero(D.SOL)
        res = dev.solve(precision=D.FP64)
        assert res["status"] == 0
        sol = dev.download(D.SOL)
        orc = O.Oracle(tree, spec)
        orc.prepare()
        L = spec.maxlevel
        tb = tower_blocks(tree, L)
        for d in range(3):
            assert np.array_equal(dev.flux(L, d, D.SOL)[tb], orc.flux(L, d, sol)[tb]), d
    finally:
        dev.close()
