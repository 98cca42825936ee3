"""Properties the mathematics fixes, checked on the C oracle independently of the golden
vectors, chosen so that a plausible slip (a dropped face, a wrong sign, a transposed index, a
wrong coarse/fine factor) fails at least one of them:

* uniform single-level grid: the operator is the textbook 7-point / 5-point -Laplacian with
  Dirichlet-by-reflection ghosts (SayakaMacOperator.cpp:1409-1454 with cell_eval_coef 1115-1167);
* composite refined tree, all-Neumann: constants are in the null space, the flux form is
  conservative and the operator is symmetric in the dv*vf inner product;
* the red-black smoother (SayakaMacOperator.cpp:1191-1323) has the exact solution as a fixed point;
* restriction of a constant and prolongation by injection (SayakaMacIB.cpp:157-283);
* BiCGSTAB (SayakaMacSolver.cpp:683-920) reaches the dense direct solve;
* the C0 manufactured solution is reproduced to second order.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

import parity_util as U
from oracle import oraclebind as O
from sayakaamr_amd import workloads as W
from sayakaamr_amd.tree import AmrTree


def uniform_case(ndim, nb, bc):
    spec = W.small_spec(ndim, maxlevel=2, nb=nb)
    tree = W.build_tree(AmrTree(ndim), spec, lambda t, nb_, nd, blocks: np.zeros(len(blocks), bool))
    orc = O.Oracle(tree, spec)
    orc.prepare()
    orc.set_bc([W.WALL], [bc], [0.0])
    return spec, tree, orc


def global_index(tree, spec):
    """Global (i, j, k) of every valid cell of every level-2 block, [n][ncell][3]."""
    nb, nd = spec.nb, spec.ndim
    n = tree["level"].shape[0]
    dx = tree["length"][:, :nd] / nb
    ii = np.indices((nb,) * nd)[::-1].reshape(nd, -1).T   # (i fastest)
    lo = np.rint(tree["lo"][:, :nd] / dx).astype(int)
    return lo[:, None, :] + ii[None, :, :]


def laplacian_1d(m, h, dirichlet):
    main = np.full(m, 2.0)
    if dirichlet:
        main[0] = main[-1] = 3.0
    else:
        main[0] = main[-1] = 1.0
    return sp.diags([main, -np.ones(m - 1), -np.ones(m - 1)], [0, -1, 1]) / h ** 2


@pytest.mark.parametrize("ndim,nb", [(2, 8), (3, 4)])
@pytest.mark.parametrize("bc", [W.DIRICHLET, W.NEUMANN])
def test_uniform_grid_is_textbook_laplacian(ndim, nb, bc):
    spec, tree, orc = uniform_case(ndim, nb, bc)
    lv2 = np.nonzero(tree["level"] == 2)[0]
    m = 2 * spec.ndiv[0] * nb   # cells per direction at level 2 (root 2^d, refined once)
    h = 1.0 / m
    T = laplacian_1d(m, h, bc == W.DIRICHLET)
    I = sp.identity(m)
    A = sp.kron(I, T) + sp.kron(T, I) if ndim == 2 else (
        sp.kron(sp.kron(I, I), T) + sp.kron(sp.kron(I, T), I) + sp.kron(sp.kron(T, I), I))
    rng = np.random.default_rng(1)
    n = tree["level"].shape[0]
    gidx = global_index(tree, spec)
    xg = rng.standard_normal((m,) * ndim)            # [k][j][i] or [j][i]
    x = np.zeros((n, nb ** ndim))
    for b in lv2:
        gi = gidx[b]
        x[b] = xg[tuple(gi[:, ::-1].T)]
    orc.buf_set(O.P, U.valid_to_grown(x, nb, ndim))
    orc.apply(2, O.V, O.P)
    out = U.grown_to_valid(orc.buf_get(O.V), nb, ndim)
    ref = (A @ xg.ravel()).reshape(xg.shape)
    got = np.zeros_like(ref)
    for b in lv2:
        got[tuple(gidx[b][:, ::-1].T)] = out[b]
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


def refined_neumann_case(with_coefs):
    spec = W.small_spec(3, 3, nb=4)
    tree = W.build_tree(AmrTree(3), spec, W.shell_criterion(r0=0.25, factor=0.25))
    orc = O.Oracle(tree, spec)
    orc.prepare()
    inp = None
    if with_coefs:
        inp = W.c1_inputs(tree, spec, seed=3)
        orc.set_ib_frac(inp["vf"], inp["af"])
        orc.set_coefb(inp["b"])
    orc.set_bc([W.WALL], [W.NEUMANN], [0.0])
    return spec, tree, orc, inp


def apply_leaves(orc, spec, tree, x):
    nb, nd, L = spec.nb, spec.ndim, int(tree["level"].max())
    orc.buf_set(O.P, U.valid_to_grown(x, nb, nd))
    orc.apply(L, O.V, O.P)
    return U.grown_to_valid(orc.buf_get(O.V), nb, nd)


def test_constants_in_null_space():
    spec, tree, orc, _ = refined_neumann_case(False)
    n = tree["level"].shape[0]
    leaf = tree["node_type"] == 1
    out = apply_leaves(orc, spec, tree, np.full((n, spec.nb ** 3), 3.25))
    hmin = tree["length"][leaf, 0].min() / spec.nb
    assert np.max(np.abs(out[leaf])) <= 1e-11 * 3.25 / hmin ** 2


@pytest.mark.parametrize("with_coefs", [False, True])
def test_conservative_and_symmetric(with_coefs):
    spec, tree, orc, inp = refined_neumann_case(with_coefs)
    n = tree["level"].shape[0]
    leaf = tree["node_type"] == 1
    vf, _, _ = orc.get_coefs()
    dv = (tree["length"][:, 0] / spec.nb) ** 3
    D = (dv[:, None] * vf)[leaf]
    rng = np.random.default_rng(5)
    x = rng.standard_normal((n, spec.nb ** 3))
    y = rng.standard_normal((n, spec.nb ** 3))
    Ax = apply_leaves(orc, spec, tree, x)[leaf]
    Ay = apply_leaves(orc, spec, tree, y)[leaf]
    scale = np.sum(np.abs(D * Ax))
    # sum of all fluxes out of all fluid cells is zero with Neumann walls
    assert abs(np.sum(D * Ax)) <= 1e-12 * scale
    # <y, D A x> == <x, D A y>
    lhs, rhs = np.sum(y[leaf] * D * Ax), np.sum(x[leaf] * D * Ay)
    assert abs(lhs - rhs) <= 1e-12 * np.sum(np.abs(y[leaf] * D * Ax))


def test_sign_convention_positive_diagonal():
    """-Laplacian: a unit spike gives a positive centre value and non-positive neighbours."""
    spec, tree, orc = uniform_case(3, 4, W.DIRICHLET)
    n = tree["level"].shape[0]
    b = int(np.nonzero(tree["level"] == 2)[0][0])
    x = np.zeros((n, 64))
    c = 1 + 4 * (1 + 4 * 1)
    x[b, c] = 1.0
    orc.buf_set(O.P, U.valid_to_grown(x, 4, 3))
    orc.apply(2, O.V, O.P)
    out = U.grown_to_valid(orc.buf_get(O.V), 4, 3)
    assert out[b, c] > 0
    out[b, c] = 0.0
    assert out.max() <= 0.0 and np.count_nonzero(out) == 6


@pytest.mark.parametrize("case", ["uniform_dirichlet", "refined_neumann_ib"])
def test_smoother_fixed_point(case):
    if case == "uniform_dirichlet":
        spec, tree, orc = uniform_case(3, 4, W.DIRICHLET)
        fluid = np.ones((tree["level"].shape[0], 64), bool)
    else:
        spec, tree, orc, _ = refined_neumann_case(True)
        vf, _, _ = orc.get_coefs()
        fluid = vf > 0
    nb, nd, L = spec.nb, spec.ndim, int(tree["level"].max())
    n = tree["level"].shape[0]
    rng = np.random.default_rng(2)
    x = rng.standard_normal((n, nb ** nd)) * fluid
    orc.buf_set(O.P, U.valid_to_grown(x, nb, nd))
    orc.apply(L, O.V, O.P)
    b = U.grown_to_valid(orc.buf_get(O.V), nb, nd)
    orc.buf_set(O.CORR, U.valid_to_grown(x, nb, nd))
    orc.buf_set(O.RESID, U.valid_to_grown(b, nb, nd))
    orc.smooth(L, O.CORR, O.RESID)
    xs = U.grown_to_valid(orc.buf_get(O.CORR), nb, nd)
    tb = U.tower_blocks(tree, L)
    assert np.max(np.abs(xs[tb] - x[tb])) <= 1e-10 * np.max(np.abs(x))


def test_restrict_constant_and_prolong_injection():
    spec, tree, orc = uniform_case(3, 4, W.NEUMANN)
    n, nb = tree["level"].shape[0], 4
    x = np.full((n, 64), 0.625)   # dyadic: the 8-term sum and the /8 are exact
    x[tree["level"] == 1] = -5.0
    orc.buf_set(O.RESID, U.valid_to_grown(x, nb, 3))
    orc.restrict(1, O.RESID)
    r = U.grown_to_valid(orc.buf_get(O.RESID), nb, 3)
    assert np.all(r[tree["level"] == 1] == 0.625)
    # prolong: child cell (i,j,k) <- parent cell ((i + off_i*nb)/2, ...) of its octant
    rng = np.random.default_rng(4)
    y = rng.standard_normal((n, 64))
    orc.buf_set(O.CORR, U.valid_to_grown(y, nb, 3))
    orc.prolong(2, O.CORR)
    p = U.grown_to_valid(orc.buf_get(O.CORR), nb, 3)
    for b in np.nonzero(tree["level"] == 2)[0]:
        par, wc = tree["parent"][b], tree["which_child"][b]
        off = [(wc >> d) & 1 for d in range(3)]
        for c in range(64):
            i, j, k = c % 4, (c // 4) % 4, c // 16
            pc = (i // 2 + 2 * off[0]) + 4 * ((j // 2 + 2 * off[1]) + 4 * (k // 2 + 2 * off[2]))
            assert p[b, c] == y[par, pc]


def test_bicgstab_reaches_dense_solve():
    spec = W.small_spec(2, 3, nb=4)
    tree = W.build_tree(AmrTree(2), spec, W.ring_criterion(r0=0.3, half=0.1))
    orc = O.Oracle(tree, spec)
    orc.prepare()
    orc.set_bc([W.WALL], [W.DIRICHLET], [0.0])
    n, nc = tree["level"].shape[0], 16
    leaf = np.nonzero(tree["node_type"] == 1)[0]
    L = int(tree["level"].max())
    # dense matrix of the composite leaf operator, column by column
    idx = [(b, c) for b in leaf for c in range(nc)]
    A = np.zeros((len(idx), len(idx)))
    for col, (b, c) in enumerate(idx):
        e = np.zeros((n, nc))
        e[b, c] = 1.0
        out = apply_leaves(orc, spec, tree, e)
        A[:, col] = out[leaf].ravel()
    rng = np.random.default_rng(9)
    rhs = np.zeros((n, nc))
    rhs[leaf] = rng.standard_normal((leaf.size, nc))
    ref = np.linalg.solve(A, rhs[leaf].ravel())
    orc.set_rhs(rhs)
    orc.set_zero_sol()
    orc.set_params(1e-12, 0.0, 2000, 1)
    res = orc.cg_solve()
    assert res["converged"]
    sol = orc.get_sol()[leaf].ravel()
    assert np.linalg.norm(sol - ref) <= 1e-9 * np.linalg.norm(ref)
    assert L == 3


def test_c0_manufactured_solution():
    spec, t, tree, inp = U.c0_case()
    orc = O.Oracle(tree, spec)
    U.setup_oracle(orc, inp)
    orc.set_rhs(inp["rhs"])
    orc.set_zero_sol()
    orc.set_params(1e-10, 0.0, 2000, 1)
    res = orc.cg_solve()
    assert res["converged"]
    leaf = tree["node_type"] == 1
    err = np.max(np.abs(orc.get_sol()[leaf] - inp["exact"][leaf]))
    # coarsest leaves h = 1/64, injection at coarse/fine faces: a few 1e-3 at most
    assert err < 5e-3, err


def _uniform_c0_error(level):
    import dataclasses
    spec = dataclasses.replace(W.C0_SPEC, maxlevel=level)
    tree = W.build_tree(AmrTree(2), spec, lambda t, nb, nd, blocks: np.ones(len(blocks), bool))
    rhs, exact = W.c0_rhs_exact(tree, spec)
    orc = O.Oracle(tree, spec)
    orc.prepare()
    orc.set_bc([W.WALL], [W.DIRICHLET], [0.0])
    orc.set_rhs(rhs)
    orc.set_zero_sol()
    orc.set_params(1e-12, 0.0, 2000, 1)
    assert orc.cg_solve()["converged"]
    leaf = tree["node_type"] == 1
    return np.max(np.abs(orc.get_sol()[leaf] - exact[leaf]))


def test_second_order_on_uniform_grids():
    e64, e128 = _uniform_c0_error(2), _uniform_c0_error(3)
    assert 3.5 < e64 / e128 < 4.5, (e64, e128)
