"""The HIP solver on the GPU, through the C ABI, against the reference's golden vectors and the
C oracle: every operator bit-exact, the plain FP64 solve bit-exact (same iterations, same
residual history, same solution)."""
from __future__ import annotations

import numpy as np
import pytest

import parity_util as U
from golden_cases import CASES, inputs_of, tower_blocks, tree_of
from oracle import oraclebind as O
from sayakaamr_amd import device as D
from sayakaamr_amd import workloads as W

pytestmark = pytest.mark.gpu


def make_device(g, spec, bc):
    tree = tree_of(g)
    inp = inputs_of(g, spec.ndim)
    n = tree["level"].shape[0]
    print(f"[gpu] device context: ndim {spec.ndim} nb {spec.nb} blocks {n}", flush=True)
    dev = D.MacSolverGPU(spec.ndim, spec.nb, n + 8, device=0)
    dev.prepare(tree, spec)
    dev.set_bc([W.WALL], [bc])
    dev.set_coefs(inp["vf"], inp["af"], inp["b"])
    return dev, tree, inp


@pytest.fixture(scope="module", params=sorted(CASES))
def dcase(request, golden):
    spec, bc = CASES[request.param]
    g = golden(request.param)
    dev, tree, inp = make_device(g, spec, bc)
    yield request.param, spec, g, dev, tree, inp, bc
    dev.close()


def test_coefficients(dcase):
    _, spec, g, dev, tree, inp, bc = dcase
    vf, b, a, w, wcf = dev.get_coefs()
    assert np.array_equal(vf, g["coef_vf"])
    for d in range(spec.ndim):
        assert np.array_equal(b[d], g[f"coef_b{d}"])
        assert np.array_equal(a[d], g[f"coef_a{d}"])


def test_apply_every_tower(dcase):
    _, spec, g, dev, tree, inp, bc = dcase
    for m in range(1, spec.maxlevel + 1):
        dev.upload(D.P, g["op_x"])
        dev.apply(m, D.V, D.P)
        tb = tower_blocks(tree, m)
        assert np.array_equal(dev.download(D.V)[tb], g[f"op_apply_m{m}"][tb]), m


def test_residual_matches_apply(dcase):
    _, spec, g, dev, tree, inp, bc = dcase
    L = spec.maxlevel
    dev.upload(D.P, g["op_x"])
    dev.upload(D.RHS, g["op_y"])
    dev.residual(L, D.R, D.P, D.RHS)
    tb = tower_blocks(tree, L)
    vf = g["coef_vf"]
    ref = np.where(vf > 0, g["op_y"] - g[f"op_apply_m{L}"], 0.0)
    assert np.array_equal(dev.download(D.R)[tb], ref[tb])


def test_smoother(dcase):
    _, spec, g, dev, tree, inp, bc = dcase
    for m in range(2, spec.maxlevel + 1):
        dev.upload(D.CORR, g["op_x"])
        dev.upload(D.RESID, g["op_y"])
        dev.smooth(m, D.CORR, D.RESID)
        tb = tower_blocks(tree, m)
        assert np.array_equal(dev.download(D.CORR)[tb], g[f"op_smooth_m{m}"][tb]), m


def test_restrict_prolong(dcase):
    _, spec, g, dev, tree, inp, bc = dcase
    L = spec.maxlevel
    dev.upload(D.RESID, g["op_y"])
    for cl in range(L - 1, 0, -1):
        dev.restrict(cl, D.RESID)
    assert np.array_equal(dev.download(D.RESID), g["op_restrict"])
    dev.upload(D.CORR, g["op_x"])
    for fl in range(2, L + 1):
        dev.prolong(fl, D.CORR)
    assert np.array_equal(dev.download(D.CORR), g["op_prolong"])


def test_dot_norm_reference_order(dcase):
    _, spec, g, dev, tree, inp, bc = dcase
    dev.upload(D.P, g["op_x"])
    dev.upload(D.V, g["op_y"])
    dot, mx = dev.dot(spec.maxlevel, D.P, D.V)
    assert dot == g["op_dot"][0] and mx == g["op_dot"][1]


def test_preconditioner(dcase):
    _, spec, g, dev, tree, inp, bc = dcase
    L = spec.maxlevel
    dev.upload(D.P, inp["rhs"])
    dev.precond(L, D.PHAT, D.P)
    tb = tower_blocks(tree, L)
    assert np.array_equal(dev.download(D.PHAT)[tb], g["op_precond"][tb])


def test_bottom_solve_matches_oracle(dcase):
    _, spec, g, dev, tree, inp, bc = dcase
    nb, nd, L = spec.nb, spec.ndim, spec.maxlevel
    orc = O.Oracle(tree, spec)
    orc.prepare()
    if inp["vf"] is not None:
        orc.set_ib_frac(inp["vf"], inp["af"])
    if inp["b"] is not None:
        orc.set_coefb(inp["b"])
    orc.set_bc([W.WALL], [bc], [0.0])
    orc.buf_set(O.RESID, U.valid_to_grown(inp["rhs"], nb, nd))
    dev.upload(D.RESID, inp["rhs"])
    for cl in range(L - 1, 0, -1):
        orc.restrict(cl, O.RESID)
        dev.restrict(cl, D.RESID)
    orc.buf_set(O.CORR, np.zeros((tree["level"].shape[0], orc.ngrown)))
    ro = orc.cg_level(1, O.CORR, O.RESID, 1e-6, 1e-6, 2000, 0)
    it, st = dev.bottom(D.CORR, D.RESID)
    t1 = tower_blocks(tree, 1)
    assert it == ro["iters"] and st == 0
    assert np.array_equal(dev.download(D.CORR)[t1], U.grown_to_valid(orc.buf_get(O.CORR), nb, nd)[t1])


def test_fp64_solve_bit_exact(dcase):
    _, spec, g, dev, tree, inp, bc = dcase
    dev.upload(D.RHS, inp["rhs"])
    dev.zero(D.SOL)
    res = dev.solve(tol_rel=1e-10, tol_abs=0.0, max_iter=2000, precond=1, precision=D.FP64)
    iters, conv, stat = g["solve_iters"]
    assert res["status"] == 0 and res["iters"] == iters
    assert np.allclose(res["hist"], g["solve_hist"], rtol=1e-6, atol=0.0)
    leaf = tree["node_type"] == 1
    assert np.array_equal(dev.download(D.SOL)[leaf], g["solve_sol"][leaf])
    # deterministic: a second solve repeats every bit
    dev.zero(D.SOL)
    res2 = dev.solve(precision=D.FP64)
    assert res2["iters"] == res["iters"] and np.array_equal(res2["hist"], res["hist"])


def test_fp64_solve_stopping_at_a_half_step(dcase):
    """A solve that stops at the half step of an iteration (MacSolver::cg_solve_level: x +=
    alpha phat, then |s| under the tolerance, SayakaMacSolver.cpp:683-920).  The device queues the
    second half of each iteration before the host reads |s|, and tests |s| on the device to skip
    the coarsest-level solve of a V-cycle whose result is not wanted; the absolute tolerance is set
    between the chosen |s| and every value tested before it.  Iterations, history, coarsest-level
    solve counts and the solution equal the C oracle's."""
    _, spec, g, dev, tree, inp, bc = dcase
    h = np.asarray(g["solve_hist"])
    tested = np.abs(h)
    half = [i for i in range(1, len(h)) if h[i] < 0 and tested[i] < tested[:i].min()]
    assert half, "no half step below every earlier test value"
    i = half[-1]
    tol = float(np.sqrt(tested[i] * tested[:i].min()))
    orc = O.Oracle(tree, spec)
    U.setup_oracle(orc, dict(inp, bc=[bc]))
    orc.set_rhs(inp["rhs"])
    orc.set_zero_sol()
    orc.set_params(0.0, tol, 2000, 1)
    orc.bottom_stats(reset=True)
    ro = orc.cg_solve()
    calls, its = orc.bottom_stats()
    dev.upload(D.RHS, inp["rhs"])
    leaf = tree["node_type"] == 1
    sol_o = orc.get_sol()[leaf]
    try:
        # 0: every second half queued before the host reads |s| (the stopping one included:
        # its coarsest solve skipped on the device); the default 100; negative: none
        for factor in (0.0, 100.0, -1.0):
            dev.set_half_step_defer(factor)
            dev.zero(D.SOL)
            res = dev.solve(tol_rel=0.0, tol_abs=tol, max_iter=2000, precond=1, precision=D.FP64)
            print(f"[gpu] half-step stop, defer factor {factor}: iters {res['iters']} (oracle {ro['iters']}), "
                  f"bottom calls {res['bottom_calls']} / {calls}, its {res['bottom_iters']} / {its}", flush=True)
            assert res["status"] == 0 and res["half_step"] and ro["converged"]
            assert res["iters"] == ro["iters"] == i // 2 + 1
            assert np.array_equal(res["hist"], ro["hist"])
            assert res["bottom_calls"] == calls == 2 * res["iters"] - 1 and res["bottom_iters"] == its
            assert np.array_equal(dev.download(D.SOL)[leaf], sol_o)
    finally:
        dev.set_half_step_defer(100.0)


def test_zero_rhs_returns_immediately(dcase):
    _, spec, g, dev, tree, inp, bc = dcase
    dev.zero(D.RHS)
    dev.zero(D.SOL)
    res = dev.solve(precision=D.FP64)
    assert res["status"] == 0 and res["iters"] == 0 and res["rnorm0"] == 0.0
    assert not np.any(dev.download(D.SOL))


def test_argument_errors_do_not_abort():
    dev = D.MacSolverGPU(3, 8, 16)
    try:
        with pytest.raises(D.DeviceError):
            dev.apply(1, 99, 0)                       # bad field id
        spec = W.small_spec(3, 3)
        from sayakaamr_amd.tree import AmrTree
        tree = W.build_tree(AmrTree(3), spec, W.shell_criterion())
        with pytest.raises(D.DeviceError):
            dev.prepare(tree, spec)                   # more blocks than the context holds
    finally:
        dev.close()
    with pytest.raises(D.DeviceError):
        D.MacSolverGPU(3, 12, 16)                     # unsupported block size
