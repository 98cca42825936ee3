"""fixBC, valued walls and warm-started solves on the GPU, bit for bit against the reference.

The device fixBC (k_fix_bc through sayaka_dev_fix_bc, and through the host MacSolver::fixBC of
the flat C ABI) against the reference's corrected rhs (tests/golden/{c0,s3}_fixbc.npz:
Dirichlet walls with values 10 / -3 / 0.5 / -2, SimpleFill walls 2.5 / 1.5, Neumann walls, each
wall its own physical code); the FP64 solve of that rhs with those walls in the operator from a
zero guess and from a non-zero setInitSol guess (SayakaMacSolver.cpp:722-735, 906-914): same
iterations, residual history and solution as the reference."""
from __future__ import annotations

import numpy as np
import pytest

import golden_cases as GC
from sayakaamr_amd import device as D
from sayakaamr_amd import solver as S
from sayakaamr_amd import workloads as W
from sayakaamr_amd.tree import AmrTree

pytestmark = pytest.mark.gpu


def _tree(name, golden):
    g = golden(name)
    spec = GC.fixbc_spec(name)
    t = AmrTree(spec.ndim, max_blocks=4000)
    tree = W.build_tree(t, spec, GC.fixbc_criterion(name))
    assert np.array_equal(tree["neighbor"], g["tree_neighbor"])
    return g, spec, t, tree, GC.fixbc_inputs(name, tree, spec)


@pytest.mark.parametrize("name", sorted(GC.FIXBC_CASES))
def test_device_fix_bc_and_valued_wall_solves(name, golden):
    g, spec, _, tree, inp = _tree(name, golden)
    n = tree["level"].shape[0]
    print(f"[gpu] fixBC + valued-wall solves, {name}: {n} blocks", flush=True)
    dev = D.MacSolverGPU(spec.ndim, spec.nb, n + 16)
    try:
        dev.prepare(tree, spec)
        codes, mb, _ = GC.fixbc_walls(name)
        dev.set_bc(codes, mb)
        dev.set_coefs(inp.get("vf"), inp.get("af"), inp.get("b"))
        dev.upload(D.RHS, inp["rhs"])
        dev.fix_bc(GC.fixbc_phi(name, tree), GC.FIXBC_CASES[name][4], D.RHS)
        assert np.array_equal(dev.download(D.RHS), g["fix_rhs"])
        dev.zero(D.SOL)
        r = dev.solve(tol_rel=1e-10, tol_abs=0.0, max_iter=2000, precond=1, precision=D.FP64)
        assert r["status"] == 0 and r["iters"] == g["solve_iters"][0]
        assert np.allclose(r["hist"], g["solve_hist"], rtol=1e-6, atol=0.0)
        assert np.array_equal(dev.download(D.SOL), g["solve_sol"])
        dev.upload(D.SOL, GC.fixbc_guess(name, tree))
        r = dev.solve(tol_rel=1e-10, tol_abs=0.0, max_iter=2000, precond=1, precision=D.FP64)
        assert r["status"] == 0 and r["iters"] == g["warm_iters"][0]
        assert np.allclose(r["hist"], g["warm_hist"], rtol=1e-6, atol=0.0)
        assert np.array_equal(dev.download(D.SOL), g["warm_sol"])
    finally:
        dev.close()


@pytest.mark.parametrize("name", sorted(GC.FIXBC_CASES))
def test_host_macsolver_fix_bc_and_warm_start(name, golden):
    g, spec, t, tree, inp = _tree(name, golden)
    print(f"[gpu] host MacSolver fixBC / setInitSol, {name}", flush=True)
    mac = S.MacSolver(t)
    try:
        mac.prepareSolver()
        codes, mb, _ = GC.fixbc_walls(name)
        mac.setBC(codes, mb)
        if inp.get("vf") is not None:
            mac.setIBFrac(inp["vf"], inp["af"])
        if inp.get("b") is not None:
            mac.setCoefB(inp["b"])
        rhs = mac.fixBC(GC.fixbc_phi(name, tree), GC.FIXBC_CASES[name][4], inp["rhs"])
        assert np.array_equal(rhs, g["fix_rhs"])
        mac.setRhs(rhs)
        mac.setInitSol(GC.fixbc_guess(name, tree))
        mac.setTolerance(1e-10, 0.0)
        mac.setMaxIteration(2000)
        assert mac.cg_solve() == 0
        assert mac.stats["status"] == 0 and mac.stats["iters"] == g["warm_iters"][0]
        assert np.array_equal(mac.getSol(), g["warm_sol"])
        with pytest.raises(S.MacSolverError):   # no ghost layer: the reference exits
            mac.fixBC(GC.fixbc_phi(name, tree)[:, :1], 0, inp["rhs"])
    finally:
        mac.close()
