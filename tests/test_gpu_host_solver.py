"""The host C++ MacSolver (include/sayaka/mac_solver.h) -- the reference's call surface -- through
its flat C ABI (include/sayaka_host.h) and as a compiled C++ driver, on the GPU."""
from __future__ import annotations

import numpy as np
import pytest

import parity_util as U
from golden_cases import tree_of
from oracle import oraclebind as O
from sayakaamr_amd import _native as N
from sayakaamr_amd import solver as S
from sayakaamr_amd import workloads as W
from sayakaamr_amd.tree import AmrTree

pytestmark = pytest.mark.gpu


def test_c1_through_host_macsolver(golden):
    g = golden("c1_summary")
    spec = W.C1_SPEC
    tree = AmrTree(3, max_blocks=6000)
    t = W.build_tree(tree, spec, W.shell_criterion())
    assert np.array_equal(t["neighbor"], g["tree_neighbor"])
    inp = W.c1_inputs(t, spec, seed=0)
    print("[gpu] host MacSolver on the 3D octree", flush=True)
    mac = S.MacSolver(tree)
    mac.prepareSolver()
    mac.setBC([W.WALL], [W.NEUMANN])
    mac.setIBFrac(inp["vf"], inp["af"])
    mac.setCoefB(inp["b"])
    mac.setRhs(inp["rhs"])
    mac.setZeroSol()
    mac.setTolerance(1e-10, 0.0)
    mac.setMaxIteration(2000)
    assert mac.cg_solve() == 0
    assert mac.stats["status"] == 0 and mac.stats["iters"] == g["solve_iters"][0]
    sol = mac.getSol()
    assert np.array_equal(sol[g["sample_blocks"], g["sample_cells"]], g["sample_sol"])
    mac.close()


def test_c0_dirichlet_through_host_macsolver(golden):
    g = golden("c0_ops")
    spec = W.C0_SPEC
    tree = AmrTree(2, max_blocks=1000)
    t = W.build_tree(tree, spec, W.ring_criterion())
    mac = S.MacSolver(tree)
    mac.prepareSolver()
    mac.setBC([W.WALL], [W.DIRICHLET])
    mac.setRhs(g["in_rhs"])
    mac.setZeroSol()
    assert mac.cg_solve() == 0 and mac.stats["iters"] == g["solve_iters"][0]
    leaf = t["node_type"] == 1
    assert np.array_equal(mac.getSol()[leaf], g["solve_sol"][leaf])
    mac.close()


def test_ib_by_level_set_matches_oracle():
    """setIBFracByLS (binary fractions from a level set, SayakaMacIB.cpp:12-154)."""
    spec = W.small_spec(3, 3, nb=8)
    tree = AmrTree(3, max_blocks=2000)
    t = W.build_tree(tree, spec, W.shell_criterion(r0=0.25, factor=0.5))
    n, nb, ng = t["level"].shape[0], 8, 1
    # level set of a sphere of radius 0.23 on the grown boxes
    ii = np.arange(-ng, nb + ng) + 0.5
    dx = t["length"][:, 0] / nb
    x = t["lo"][:, 0, None] + ii[None, :] * dx[:, None]
    y = t["lo"][:, 1, None] + ii[None, :] * dx[:, None]
    z = t["lo"][:, 2, None] + ii[None, :] * dx[:, None]
    ls = np.sqrt((x[:, None, None, :] - 0.5) ** 2 + (y[:, None, :, None] - 0.5) ** 2 +
                 (z[:, :, None, None] - 0.5) ** 2) - 0.23
    ls = ls.reshape(n, -1)
    rng = np.random.default_rng(6)
    rhs = rng.standard_normal((n, nb ** 3))
    orc = O.Oracle(t, spec)
    orc.prepare()
    orc.set_ib_ls(ls, ng)
    orc.set_bc([W.WALL], [W.NEUMANN], [0.0])
    vf, _, _ = orc.get_coefs()
    rhs = np.where(vf > 0, rhs, 0.0)
    leaf = t["node_type"] == 1
    dv = dx ** 3
    rhs[leaf] -= np.sum((dv[:, None] * vf * rhs)[leaf]) / np.sum((dv[:, None] * vf)[leaf]) * (vf[leaf] > 0)
    orc.set_rhs(rhs)
    orc.set_zero_sol()
    orc.set_params(1e-10, 0.0, 2000, 1)
    ro = orc.cg_solve()
    mac = S.MacSolver(tree)
    mac.prepareSolver()
    mac.setBC([W.WALL], [W.NEUMANN])
    mac.setIBFracByLS(ls, ng)
    mac.setRhs(rhs)
    mac.setZeroSol()
    assert mac.cg_solve() == 0
    assert mac.stats["iters"] == ro["iters"]
    assert np.array_equal(mac.getSol()[leaf], orc.get_sol()[leaf])
    mac.close()
