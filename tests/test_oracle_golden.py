"""The C oracle (oracle/sayaka_oracle.c) against the reference's golden vectors: bit-exact.

Each operator of the reference solve path -- coefficients (SayakaMacCoef.cpp, SayakaMacIB.cpp),
apply (SayakaMacOperator.cpp:84-137), red-black smoother (140-159), restriction / prolongation
(SayakaMacIB.cpp:157-283), dot / max-norm (SayakaMacUtil.cpp:53-106), the V-cycle
preconditioner (SayakaMacSolver.cpp:245-374, 924-949) and the full MG-BiCGSTAB solve
(455-920) -- is run on the recorded inputs and compared bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

import parity_util as U
from golden_cases import CASES, inputs_of, tower_blocks, tree_of
from oracle import oraclebind as O
from sayakaamr_amd import workloads as W


def make_oracle(g, spec, bc):
    tree = tree_of(g)
    inp = inputs_of(g, spec.ndim)
    orc = O.Oracle(tree, spec)
    orc.prepare()
    if inp["vf"] is not None:
        orc.set_ib_frac(inp["vf"], inp["af"])
    if inp["b"] is not None:
        orc.set_coefb(inp["b"])
    orc.set_bc([W.WALL], [bc], [0.0])
    return orc, tree, inp


@pytest.fixture(params=sorted(CASES))
def case(request, golden):
    spec, bc = CASES[request.param]
    g = golden(request.param)
    orc, tree, inp = make_oracle(g, spec, bc)
    return request.param, spec, g, orc, tree, inp


def test_coefficients(case):
    _, spec, g, orc, tree, inp = case
    vf, b, a = orc.get_coefs()
    assert np.array_equal(vf, g["coef_vf"])
    for d in range(spec.ndim):
        assert np.array_equal(b[d], g[f"coef_b{d}"]), d
        assert np.array_equal(a[d], g[f"coef_a{d}"]), d


def test_towers(case):
    _, spec, g, orc, tree, inp = case
    for m in range(1, spec.maxlevel + 1):
        tw = orc.tower(m)
        for k, v in tw.items():
            assert np.array_equal(v, g[f"tower_m{m}_{k}"]), (m, k)


def test_apply_every_tower(case):
    _, spec, g, orc, tree, inp = case
    nb, nd = spec.nb, spec.ndim
    for m in range(1, spec.maxlevel + 1):
        orc.buf_set(O.P, U.valid_to_grown(g["op_x"], nb, nd))
        orc.apply(m, O.V, O.P)
        tb = tower_blocks(tree, m)
        out = U.grown_to_valid(orc.buf_get(O.V), nb, nd)
        assert np.array_equal(out[tb], g[f"op_apply_m{m}"][tb]), m


def test_smoother(case):
    _, spec, g, orc, tree, inp = case
    nb, nd = spec.nb, spec.ndim
    for m in range(2, spec.maxlevel + 1):
        orc.buf_set(O.CORR, U.valid_to_grown(g["op_x"], nb, nd))
        orc.buf_set(O.RESID, U.valid_to_grown(g["op_y"], nb, nd))
        orc.smooth(m, O.CORR, O.RESID)
        tb = tower_blocks(tree, m)
        out = U.grown_to_valid(orc.buf_get(O.CORR), nb, nd)
        assert np.array_equal(out[tb], g[f"op_smooth_m{m}"][tb]), m


def test_restrict_prolong(case):
    _, spec, g, orc, tree, inp = case
    nb, nd, L = spec.nb, spec.ndim, spec.maxlevel
    orc.buf_set(O.RESID, U.valid_to_grown(g["op_y"], nb, nd))
    for cl in range(L - 1, 0, -1):
        orc.restrict(cl, O.RESID)
    assert np.array_equal(U.grown_to_valid(orc.buf_get(O.RESID), nb, nd), g["op_restrict"])
    orc.buf_set(O.CORR, U.valid_to_grown(g["op_x"], nb, nd))
    for fl in range(2, L + 1):
        orc.prolong(fl, O.CORR)
    assert np.array_equal(U.grown_to_valid(orc.buf_get(O.CORR), nb, nd), g["op_prolong"])


def test_dot_norm(case):
    _, spec, g, orc, tree, inp = case
    nb, nd, L = spec.nb, spec.ndim, spec.maxlevel
    orc.buf_set(O.P, U.valid_to_grown(g["op_x"], nb, nd))
    orc.buf_set(O.V, U.valid_to_grown(g["op_y"], nb, nd))
    assert orc.dot(L, O.P, O.V) == g["op_dot"][0]
    assert orc.norm(L, O.P) == g["op_dot"][1]


def test_preconditioner(case):
    _, spec, g, orc, tree, inp = case
    nb, nd, L = spec.nb, spec.ndim, spec.maxlevel
    orc.buf_set(O.P, U.valid_to_grown(inp["rhs"], nb, nd))
    orc.precond(L, O.PHAT, O.P)
    tb = tower_blocks(tree, L)
    out = U.grown_to_valid(orc.buf_get(O.PHAT), nb, nd)
    assert np.array_equal(out[tb], g["op_precond"][tb])


def test_full_solve(case):
    _, spec, g, orc, tree, inp = case
    orc.set_rhs(inp["rhs"])
    orc.set_zero_sol()
    orc.set_params(1e-10, 0.0, 2000, 1)
    res = orc.cg_solve()
    iters, conv, stat = g["solve_iters"]
    assert res["iters"] == iters and res["converged"] == bool(conv)
    # the reference's history comes from its log lines (printf %e: 7 significant digits)
    assert np.allclose(res["hist"], g["solve_hist"], rtol=1e-6, atol=0.0)
    leaf = tree["node_type"] == 1
    assert np.array_equal(orc.get_sol()[leaf], g["solve_sol"][leaf])


def test_c1_full_solve(golden):
    """The full-size 3D workload: iterations, residual history and sampled solution."""
    g = golden("c1_summary")
    spec = W.C1_SPEC
    tree = tree_of(g)
    inp = W.c1_inputs(tree, spec, seed=0)
    assert inp["rhs"].sum() == g["rhs_checksum"][0]
    assert np.abs(inp["rhs"]).sum() == g["rhs_checksum"][1]
    orc = O.Oracle(tree, spec)
    orc.prepare()
    orc.set_ib_frac(inp["vf"], inp["af"])
    orc.set_coefb(inp["b"])
    orc.set_bc([W.WALL], [W.NEUMANN], [0.0])
    orc.set_rhs(inp["rhs"])
    orc.set_zero_sol()
    orc.set_params(1e-10, 0.0, 2000, 1)
    res = orc.cg_solve()
    assert res["iters"] == g["solve_iters"][0]
    assert np.allclose(res["hist"], g["solve_hist"], rtol=1e-6, atol=0.0)
    sol = orc.get_sol()
    assert np.array_equal(sol[g["sample_blocks"], g["sample_cells"]], g["sample_sol"])
