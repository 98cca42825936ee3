"""The C oracle's face flux (MacSolver::mg_flux_level, SayakaMacOperator.cpp:161-239, 929-1025)
against the reference's golden vectors, bit for bit, and against closed forms.

Golden: tests/golden/{c0,s3}_flux.npz, written by tests/golden/make_golden.py from the
compiled reference: the flux of every direction on the tower of every MG level for a seeded
field, and on the finest tower for the case's own solution.

Closed form: for phi = a x (at cell centres) the x-flux is -a on every face that is not a wall
-- same-level faces (-(phi_c - phi_w)/dx), the fine side of a coarse/fine face (the centres are
1.5 fine dx apart, which the reference's 1.5 dh divisor accounts for) and the coarse side (the
average of the fine fluxes) -- and 0 on walls.  A wrong divisor, sign or averaging weight fails
it.
"""
from __future__ import annotations

import numpy as np
import pytest

from golden_cases import CASES, inputs_of, tower_blocks, tree_of
from oracle import oraclebind as O
from sayakaamr_amd import workloads as W
from sayakaamr_amd.tree import AmrTree

FLUX_FILES = {"c0_ops": "c0_flux", "s3_ops": "s3_flux"}


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
    return orc, tree


@pytest.fixture(params=sorted(CASES))
def fcase(request, golden):
    spec, bc = CASES[request.param]
    g = golden(request.param)
    gf = golden(FLUX_FILES[request.param])
    orc, tree = make_oracle(g, spec, bc)
    return spec, g, gf, orc, tree


def test_flux_every_tower_bit_exact(fcase):
    spec, g, gf, orc, tree = fcase
    assert np.array_equal(gf["op_x"], g["op_x"])
    for m in range(1, spec.maxlevel + 1):
        for d in range(spec.ndim):
            f = orc.flux(m, d, gf["op_x"])
            assert np.array_equal(f, gf[f"flux_m{m}_d{d}"]), (m, d)


def test_flux_of_solution_bit_exact(fcase):
    spec, g, gf, orc, tree = fcase
    for d in range(spec.ndim):
        assert np.array_equal(orc.flux(spec.maxlevel, d, gf["sol"]), gf[f"flux_sol_d{d}"]), d


@pytest.mark.parametrize("ndim", [2, 3])
def test_flux_of_linear_field_is_minus_slope(ndim):
    spec = W.small_spec(ndim, 3, nb=4 if ndim == 3 else 8)
    crit = W.shell_criterion(r0=0.25, factor=0.25) if ndim == 3 else W.ring_criterion()
    tree = W.build_tree(AmrTree(ndim), spec, crit)
    orc = O.Oracle(tree, spec)
    orc.prepare()
    orc.set_bc([W.WALL], [W.NEUMANN], [0.0])
    x, y, z = W.cell_centres(tree, spec.nb, ndim)
    n = tree["level"].shape[0]
    slope = 1.75
    for d, coord in enumerate((x, y, z)[:ndim]):
        phi = (slope * coord).reshape(n, -1)
        for m in range(2, spec.maxlevel + 1):
            f = orc.flux(m, d, phi)
            tb = tower_blocks(tree, m)
            fd = (spec.nb + 1) * spec.nb ** (ndim - 1)
            shape = [spec.nb] * ndim
            shape[d] += 1
            for b in tb:
                fb = f[b].reshape(shape[::-1])                     # [k][j][i] / [j][i]
                ax = ndim - 1 - d                                  # axis of direction d
                lo_wall = tree["neighbor"][b, 2 * d] <= -20
                hi_wall = tree["neighbor"][b, 2 * d + 1] <= -20
                inner = np.take(fb, range(1, spec.nb), axis=ax)
                np.testing.assert_allclose(inner, -slope, rtol=1e-12)
                lo = np.take(fb, 0, axis=ax)
                hi = np.take(fb, spec.nb, axis=ax)
                if lo_wall:
                    assert np.all(lo == 0.0)
                else:
                    np.testing.assert_allclose(lo, -slope, rtol=1e-12)
                if hi_wall:
                    assert np.all(hi == 0.0)
                else:
                    np.testing.assert_allclose(hi, -slope, rtol=1e-12)
                assert fb.size == fd


@pytest.mark.parametrize("ndim", [2, 3])
def test_projection_with_oracle_flux_is_divergence_free(ndim):
    """Solve L p = div(u) with unit coefficients, correct u -= flux(p): the corrected velocity
    is divergence-free on every leaf cell (the operator is the divergence of the flux)."""
    spec = W.small_spec(ndim, 3, nb=4 if ndim == 3 else 8)
    crit = W.shell_criterion(r0=0.25, factor=0.25) if ndim == 3 else W.ring_criterion()
    tree = W.build_tree(AmrTree(ndim), spec, crit)
    n = tree["level"].shape[0]
    L = spec.maxlevel
    nf = (spec.nb + 1) * spec.nb ** (ndim - 1)
    one_c = np.ones((n, spec.nb ** ndim))
    one_f = [np.ones((n, nf)) for _ in range(ndim)]
    u = W.face_velocity(tree, spec, seed=5)
    rhs = W.divergence_rhs(tree, spec, u, one_c, one_f)
    orc = O.Oracle(tree, spec)
    orc.prepare()
    orc.set_bc([W.WALL], [W.NEUMANN], [0.0])
    orc.set_rhs(rhs)
    orc.set_zero_sol()
    orc.set_params(1e-12, 0.0, 2000, 1)
    res = orc.cg_solve()
    assert res["converged"]
    p = orc.get_sol()
    leaf = tree["node_type"] == 1
    div1 = W.divergence_rhs(tree, spec, [u[d] - orc.flux(L, d, p) for d in range(ndim)], one_c, one_f)
    assert np.abs(div1[leaf]).max() <= 1e-9 * np.abs(rhs[leaf]).max()
    # a flux without the coarse/fine 1.5 factor breaks it
    bad = [orc.flux(L, d, p) for d in range(ndim)]
    tb = tower_blocks(tree, L)
    shape = [spec.nb] * ndim
    for d in range(ndim):
        for b in tb:
            for s in (0, 1):
                if tree["neighbor"][b, 2 * d + s] == -1:       # coarse neighbour: fine side of a c/f face
                    sh = list(shape)
                    sh[d] += 1
                    fb = bad[d][b].reshape(sh[::-1])
                    idx = [slice(None)] * ndim
                    idx[ndim - 1 - d] = 0 if s == 0 else spec.nb
                    fb[tuple(idx)] *= 1.5
    div2 = W.divergence_rhs(tree, spec, [u[d] - bad[d] for d in range(ndim)], one_c, one_f)
    assert np.abs(div2[leaf]).max() > 1e-3 * np.abs(rhs[leaf]).max()
