"""MacSolver::fixBC, valued walls and warm-started solves: the C oracle against the reference.

fixBC / fixBlockBC (SayakaMacOperator.cpp:1470-1594) adds the wall terms of Dirichlet and
SimpleFill walls to the rhs from the BC values phi's wall ghosts carry (Test6.h:270-276,
607-609).  The golden vectors (tests/golden/{c0,s3}_fixbc.npz, written by the compiled
reference) hold: the corrected rhs on a 2D quadtree and a 3D octree whose walls each have their
own BC (Dirichlet with values 10, -3, 0.5, -2; SimpleFill 2.5 and 1.5; Neumann); the solve of
that rhs with those walls in the operator from a zero guess; and the same solve from a non-zero
setInitSol guess (SayakaMacSolver.cpp:722-735, 906-914).
"""
from __future__ import annotations

import numpy as np
import pytest

import golden_cases as GC
from oracle import oraclebind as O
from sayakaamr_amd import workloads as W
from sayakaamr_amd.tree import AmrTree


def _setup(name, golden):
    g = golden(name)
    spec = GC.fixbc_spec(name)
    t = AmrTree(spec.ndim, max_blocks=4000)
    tree = W.build_tree(t, spec, GC.fixbc_criterion(name))
    for k in W.TREE_FIELDS:
        assert np.array_equal(tree[k], g[f"tree_{k}"]), k
    inp = GC.fixbc_inputs(name, tree, spec)
    assert np.array_equal(inp["rhs"], g["in_rhs"])
    orc = O.Oracle(tree, spec)
    orc.prepare()
    if inp.get("vf") is not None:
        orc.set_ib_frac(inp["vf"], inp["af"])
    if inp.get("b") is not None:
        orc.set_coefb(inp["b"])
    orc.set_bc(*GC.fixbc_walls(name))
    return g, spec, tree, inp, orc


@pytest.mark.parametrize("name", sorted(GC.FIXBC_CASES))
def test_fix_bc_matches_reference(name, golden):
    g, spec, tree, inp, orc = _setup(name, golden)
    phi = GC.fixbc_phi(name, tree)
    out = orc.fix_bc(phi, GC.FIXBC_CASES[name][4], inp["rhs"])
    assert np.array_equal(out, g["fix_rhs"])
    assert (out != inp["rhs"]).sum() > 0


@pytest.mark.parametrize("name", sorted(GC.FIXBC_CASES))
def test_fix_bc_closed_form(name, golden):
    """On a wall cell the term is 2 w phi_bc / (dv vf), with phi_bc the wall value (Dirichlet:
    the mean of the ghost and the inside value, which the BC fill set to the value; SimpleFill:
    the ghost); w = beta ds/dx B a_f of the wall face.  The face weights come from the oracle's
    coefficient state; the rest from the geometry.  A dropped factor 2, a wrong ghost or a wrong
    face index fails this."""
    g, spec, tree, inp, orc = _setup(name, golden)
    nd, nb = spec.ndim, spec.nb
    codes, mb, bv = GC.fixbc_walls(name)
    vf, bco, af = orc.get_coefs()
    n = tree["level"].shape[0]
    lv, leaf = tree["level"], tree["node_type"] == 1
    L = int(lv.max())
    tower = ((lv < L) & leaf) | (lv == L)
    expect = inp["rhs"].copy()
    nz = nb if nd == 3 else 1
    for b in np.nonzero(tower)[0]:
        dx = tree["length"][b, 0] / nb
        dv = dx ** nd
        for f in range(2 * nd):
            code = int(tree["neighbor"][b, f])
            if code > -20:
                continue
            q = codes.index(code)
            if mb[q] == W.NEUMANN:
                continue
            d, s = f // 2, f % 2
            fx, fy = nb + (d == 0), nb + (d == 1)
            for k in range(nz):
                for j in range(nb):
                    for i in range(nb):
                        ijk = [i, j, k]
                        if ijk[d] != (nb - 1 if s else 0):
                            continue
                        fq = list(ijk)
                        fq[d] += s
                        fi = fq[0] + fx * (fq[1] + fy * fq[2])
                        w = 1.0 * (dv / dx) / dx * bco[d][b, fi] * af[d][b, fi]
                        c = i + nb * (j + nb * k)
                        if vf[b, c] > 0:
                            expect[b, c] += 2.0 * w * bv[q] / (dv * vf[b, c])
    out = g["fix_rhs"]
    scale = np.abs(expect).max()
    assert np.allclose(out, expect, rtol=1e-12, atol=1e-12 * scale)


@pytest.mark.parametrize("name", sorted(GC.FIXBC_CASES))
def test_valued_wall_solves_match_reference(name, golden):
    """The solve of the fixBC rhs (zero guess) and from a non-zero initial guess: iterations,
    residual history and solution bit for bit."""
    g, spec, tree, inp, orc = _setup(name, golden)
    orc.set_rhs(g["fix_rhs"])
    orc.set_zero_sol()
    orc.set_params(1e-10, 0.0, 2000, 1)
    r = orc.cg_solve()
    assert r["iters"] == g["solve_iters"][0] and r["stat"] == g["solve_iters"][2]
    # the reference's history comes from its log lines (printf %e: 7 significant digits)
    assert np.allclose(r["hist"], g["solve_hist"], rtol=1e-6, atol=0.0)
    assert np.array_equal(orc.get_sol(), g["solve_sol"])
    orc.set_init_sol(GC.fixbc_guess(name, tree))
    r = orc.cg_solve()
    assert r["iters"] == g["warm_iters"][0] and r["stat"] == g["warm_iters"][2]
    assert np.allclose(r["hist"], g["warm_hist"], rtol=1e-6, atol=0.0)
    assert np.array_equal(orc.get_sol(), g["warm_sol"])


@pytest.mark.parametrize("name", sorted(GC.FIXBC_CASES))
def test_warm_and_cold_solves_agree(name, golden):
    """Both solves solve the same system (walls with Dirichlet parts: no null space), so they
    agree to the solve's tolerance on the fluid leaf cells; the warm one is not the cold one bit
    for bit.  Solid cells (vf = 0) are outside the system and keep the initial guess."""
    g = golden(name)
    spec = GC.fixbc_spec(name)
    tree = {k: g[f"tree_{k}"] for k in W.TREE_FIELDS}
    inp = GC.fixbc_inputs(name, tree, spec)
    leaf = g["tree_node_type"] == 1
    fluid = np.ones_like(g["solve_sol"], bool) if inp.get("vf") is None else inp["vf"] > 0
    a, b = g["solve_sol"][leaf][fluid[leaf]], g["warm_sol"][leaf][fluid[leaf]]
    assert not np.array_equal(a, b)
    assert np.linalg.norm(a - b) <= 1e-7 * np.linalg.norm(a)
