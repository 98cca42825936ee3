"""MacSolver::mg_solve and relax_solve (SayakaMacSolver.cpp:43-162, 377-452) through the host
MacSolver on the GPU, against the reference's own solutions (tests/golden/{c0,s3}_mgrelax.npz:
C0, 2D Dirichlet; a 3D octree of 4^3 blocks with cut cells and variable B, Neumann): 3 fixed
V-cycles, V-cycles to 1e-8, and 6 relaxation passes, each from a zero solution.  Bit-exact on
every block."""
from __future__ import annotations

import numpy as np
import pytest

from sayakaamr_amd import solver as S
from sayakaamr_amd import workloads as W
from sayakaamr_amd.tree import AmrTree

pytestmark = pytest.mark.gpu

CASES = {
    "c0_mgrelax": (W.C0_SPEC, lambda: W.ring_criterion(), W.DIRICHLET),
    "s3_mgrelax": (W.small_spec(3, 3, nb=4), lambda: W.shell_criterion(r0=0.25, factor=0.25), W.NEUMANN),
}


@pytest.mark.parametrize("name", list(CASES))
def test_mg_and_relax_solve_match_reference(name, golden):
    g = golden(name)
    spec, crit, bc = CASES[name]
    tree = AmrTree(spec.ndim, max_blocks=2000)
    t = W.build_tree(tree, spec, crit())
    inp = W.c1_inputs(t, spec, seed=0) if spec.ndim == 3 else dict(rhs=W.c0_rhs_exact(t, spec)[0])
    print(f"[gpu] mg/relax solve {name}: {t['level'].shape[0]} blocks", flush=True)
    mac = S.MacSolver(tree)
    try:
        mac.prepareSolver()
        if inp.get("vf") is not None:
            mac.setIBFrac(inp["vf"], inp["af"])
        if inp.get("b") is not None:
            mac.setCoefB(inp["b"])
        mac.setBC([W.WALL], [bc])
        mac.setRhs(inp["rhs"])
        mac.setTolerance(1e-8, 0.0)
        mac.setMaxIteration(40)
        mac.setZeroSol()
        assert mac.mg_solve(3) == 0 and mac.stats["iters"] == 3
        assert np.array_equal(mac.getSol().view(np.int64), g["mg_fixed3"].view(np.int64))
        mac.setZeroSol()
        assert mac.mg_solve() == 0 and mac.stats["status"] == 0
        assert mac.stats["rnorm"] < 1e-8 * mac.stats["rnorm0"]
        assert np.array_equal(mac.getSol().view(np.int64), g["mg_tol"].view(np.int64))
        mac.setTolerance(1e-12, 0.0)
        mac.setMaxIteration(6)
        mac.setZeroSol()
        assert mac.relax_solve() == 0 and mac.stats["iters"] == 6
        assert np.array_equal(mac.getSol().view(np.int64), g["relax6"].view(np.int64))
    finally:
        mac.close()
