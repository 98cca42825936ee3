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


def test_projection_through_mg_flux_level():
    """The projection step of the reference's flow demo (Test7.h:612-653): solve for the
    pressure of rhs = div(u), take its face flux with mg_flux_level and correct u -= flux.  With
    unit coefficients the operator is exactly the divergence of that flux (in-block and
    same-level faces over dx, fine side of coarse/fine faces over 1.5 dx, coarse side the
    average of the fine fluxes), so the corrected velocity is divergence-free on every leaf cell
    of the 3D octree to the solver's residual (1e-10 relative; the reference's algorithm stalls
    before 1e-12 on this tree).  A flux with a wrong sign, divisor or coarse/fine
    average leaves an O(1) divergence at the faces where it is wrong."""
    spec = W.C1_SPEC
    tree = AmrTree(3, max_blocks=6000)
    t = W.build_tree(tree, spec, W.shell_criterion())
    n = t["level"].shape[0]
    L = int(t["level"].max())
    one_c = np.ones((n, spec.nb ** 3))
    one_f = [np.ones((n, (spec.nb + 1) * spec.nb ** 2)) for _ in range(3)]
    u = W.face_velocity(t, spec, seed=3)
    rhs = W.divergence_rhs(t, spec, u, one_c, one_f)
    mac = S.MacSolver(tree)
    try:
        mac.prepareSolver()
        mac.setBC([W.WALL], [W.NEUMANN])
        mac.setRhs(rhs)
        mac.setZeroSol()
        mac.setTolerance(1e-10, 0.0)
        assert mac.cg_solve() == 0 and mac.stats["status"] == 0
        p = mac.getSol()
        flux = [mac.mg_flux_level(L, d, p) for d in range(3)]
    finally:
        mac.close()
    leaf = t["node_type"] == 1
    div0 = np.abs(rhs[leaf]).max()
    div1 = W.divergence_rhs(t, spec, [u[d] - flux[d] for d in range(3)], one_c, one_f)
    print(f"[gpu] projection: |div u| {div0:.3e} -> {np.abs(div1[leaf]).max():.3e}", flush=True)
    # the residual of a 1e-10 solve, plus rounding of the divergence sums
    assert np.abs(div1[leaf]).max() <= 1e-8 * div0
