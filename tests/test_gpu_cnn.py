"""The learned initial guess on the GPU: input generator, CNN windows, forward output and
parameter gradients against the CPU oracle (oracle/cnn_oracle.py), training, and a solve that
starts from the learned guess (with the zero-guess fallback)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import cnn_oracle as K
from oracle import oraclebind as O
from sayakaamr_amd import device as D
from sayakaamr_amd import warmstart as WS
from sayakaamr_amd import workloads as W
from sayakaamr_amd.tree import AmrTree

pytestmark = pytest.mark.gpu

SMALL = W.small_spec(3, 3)
CENTRE, RADIUS, RATIO = (0.45, 0.52, 0.5), 0.23, 100.0
BLOBS = dict(c=np.array([[0.3, 0.4, 0.5], [0.7, 0.6, 0.5], [0.5, 0.5, 0.8]]), w=np.array([0.05, 0.08, 0.03]),
             a=np.array([1.0, -0.7, 0.4]))


@pytest.fixture(scope="module")
def small():
    tree = AmrTree(3, max_blocks=2000).build_shell(SMALL, CENTRE, RADIUS, W.SHELL_FACTOR)
    n = tree["level"].shape[0]
    print(f"[gpu] learned-guess small tree: {n} blocks", flush=True)
    dev = D.MacSolverGPU(3, 8, n + 8, device=0)
    dev.cnn_setup(64)
    dev.prepare(tree, SMALL)
    dev.set_bc([W.WALL], [W.NEUMANN])
    eps = 1.5 * (1.0 / (SMALL.ndiv[0] * 8 * 2 ** (SMALL.maxlevel - 1)))
    mean = dev.gen_sphere_blobs(tree, CENTRE, RADIUS, RATIO, eps, 4, BLOBS)
    yield dev, tree, eps, mean
    dev.close()


def test_generator_matches_workload_definitions(small):
    dev, tree, eps, mean = small
    vf, af = W.sphere_fractions(tree, SMALL, CENTRE, RADIUS)
    b = W.density_coef(tree, SMALL, CENTRE, RADIUS, RATIO, eps)
    orc = O.Oracle(tree, SMALL)
    orc.prepare()
    orc.set_ib_frac(vf, af)
    orc.set_coefb(b)
    vo, bo, ao = orc.get_coefs()
    vd, bd, ad, _, _ = dev.get_coefs()
    assert np.array_equal(vd, vo)
    for d in range(3):
        assert np.array_equal(ad[d], ao[d])
        assert np.allclose(bd[d], bo[d], rtol=1e-14, atol=0.0)
    # rhs: the blob sum on fluid cells minus its volume-weighted mean over the fluid leaves
    x, y, z = W.cell_centres(tree, 8, 3)
    f = np.zeros(x.shape)
    for c, w, a in zip(BLOBS["c"], BLOBS["w"], BLOBS["a"]):
        f += a * np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / (2 * w * w))
    f = np.where(vo > 0, f.reshape(vo.shape), 0.0)
    leaf = tree["node_type"] == 1
    dv = (tree["length"][:, 0] / 8) ** 3
    wv = (dv[:, None] * vo)[leaf]
    m = np.sum(wv * f[leaf]) / np.sum(wv)
    assert abs(m - mean) <= 1e-12 * np.abs(f).max()
    assert np.allclose(dev.download(D.RHS), np.where(vo > 0, f - m, 0.0), rtol=0, atol=1e-12 * np.abs(f).max())


def numpy_windows(tree, ids, rhs, vf, b, af, s_rhs):
    """The 12^3 windows of k_cnn_gather, restated (before BF16 rounding)."""
    out = np.zeros((len(ids), 12 ** 3, 16))
    nbr, level = tree["neighbor"], tree["level"]
    for s, blk in enumerate(ids):
        for v in range(12 ** 3):
            c3 = [v % 12 - 2, (v // 12) % 12 - 2, v // 144 - 2]
            outside = [d for d in range(3) if not 0 <= c3[d] < 8]
            src = blk
            if len(outside) == 1:
                d = outside[0]
                f = 2 * d + (c3[d] >= 8)
                n = nbr[blk, f]
                src = n if n >= 0 and level[n] == level[blk] else -1
                c3[d] += 8 if c3[d] < 0 else -8
            elif len(outside) > 1:
                src = -1
            if src < 0:
                continue
            cell = c3[0] + 8 * (c3[1] + 8 * c3[2])
            out[s, v, 0] = np.float32(rhs[src, cell] * s_rhs)
            for d in range(3):
                fx, fy = 8 + (d == 0), 8 + (d == 1)
                lo = c3[0] + fx * (c3[1] + fy * c3[2])
                hi3 = [c3[0] + (d == 0), c3[1] + (d == 1), c3[2] + (d == 2)]
                hi = hi3[0] + fx * (hi3[1] + fy * hi3[2])
                out[s, v, 1 + d] = np.float32(0.5 * (b[d][src, lo] * af[d][src, lo] + b[d][src, hi] * af[d][src, hi]))
            out[s, v, 4] = np.float32(vf[src, cell])
    return out


@pytest.fixture(scope="module")
def windows(small):
    dev, tree, eps, mean = small
    leaves = np.nonzero(tree["node_type"] == 1)[0]
    rng = np.random.default_rng(5)
    ids = rng.choice(leaves, 3, replace=False).astype(np.int32)
    dev.upload(D.SOL, rng.normal(0, 1, (tree["level"].shape[0], 512)))
    dev.cnn_store_reset(16)
    rmax = np.abs(dev.download(D.RHS)).max()
    n = dev.cnn_collect(ids, 1.0 / rmax, 0.25, 2.0)
    assert n == 3
    return ids, 1.0 / rmax


def test_windows_and_targets(small, windows):
    dev, tree, eps, mean = small
    ids, s_rhs = windows
    vf, b, af, _, _ = dev.get_coefs()
    rhs, sol = dev.download(D.RHS), dev.download(D.SOL)
    w, t = dev.cnn_store_read(0, 3)
    ref = K.bf16(numpy_windows(tree, ids, rhs, vf, b, af, s_rhs))
    assert np.array_equal(w, ref)
    assert np.array_equal(t, ((sol[ids] - 0.25) * 2.0).astype(np.float32))


def test_forward_and_gradients_match_oracle(small, windows):
    dev, tree, eps, mean = small
    rng = np.random.default_rng(7)
    p = WS.init_params(3)
    for l in range(4):
        p[l * K.PL + 27 * 256:(l + 1) * K.PL] = (rng.normal(0, 0.05, 16) * (1 if l < 3 else np.eye(16)[0])).astype(np.float32)
    dev.cnn_set_params(p)
    w, t = dev.cnn_store_read(0, 3)
    loss_d, g_d, out_d = dev.cnn_loss_grad(np.arange(3), want_out=True)
    loss_o, g_o, out_o = K.loss_and_grad(p.astype(np.float64), w.astype(np.float64), t.astype(np.float64))
    out_o = out_o.reshape(3, -1)
    # device and oracle round to BF16 at the same points; they differ by FP32 vs FP64 sums
    assert np.linalg.norm(out_d - out_o) <= 2e-3 * np.linalg.norm(out_o)
    assert abs(loss_d - loss_o) <= 2e-3 * loss_o
    for l in range(4):
        sl = slice(l * K.PL, (l + 1) * K.PL)
        assert np.linalg.norm(g_d[sl] - g_o[sl]) <= 2e-2 * np.linalg.norm(g_o[sl]), l


def test_adam_step_matches_oracle_and_training_descends(small, windows):
    dev, tree, eps, mean = small
    p0 = WS.init_params(4)
    dev.cnn_set_params(p0)
    w, t = dev.cnn_store_read(0, 3)
    _, g_o, _ = K.loss_and_grad(p0.astype(np.float64), w.astype(np.float64), t.astype(np.float64))
    dev.cnn_train_step(np.arange(3), 1e-3)
    p1 = dev.cnn_get_params()
    q, _, _ = K.adam(p0.astype(np.float64), g_o, np.zeros(p0.size), np.zeros(p0.size), 1, lr=1e-3)
    big = np.abs(g_o) > 1e-3 * np.abs(g_o).max()
    # the first Adam step moves every parameter with a clear gradient by lr against its sign
    assert np.mean(np.abs(p1[big] - q[big]) <= 1e-6) > 0.99
    assert np.all(p1[~big & (g_o == 0)] == p0[~big & (g_o == 0)])
    losses = [dev.cnn_train_step(np.arange(3), 3e-3) for _ in range(100)]
    assert losses[-1] < 0.9 * losses[0], (losses[0], losses[-1])


def test_learned_guess_solve_reaches_target():
    dev = WS.make_device(max_blocks=12000, windows=1024)
    try:
        dev.cnn_set_params(WS.init_params(0))
        pb = WS.Problem(dev, WS.sample(np.random.default_rng(123)))
        r = WS.solve_pair(dev, pb, D.MIXED)
        print("[gpu] learned guess", r, flush=True)
        assert r["status"] == (0, 0)
        # the final residual meets the absolute target whether or not the guess was used
        assert pb.residual_inf(D.SOL) <= 2e-10 * pb.bmax
    finally:
        dev.close()


def test_guess_study_on_held_out_problems():
    """The trained network and the exact-correction bounds on three held-out problems
    (warmstart.study_problem): every solve reaches its target; the fallback never costs
    iterations against the zero guess (deterministic solves of the same system); the trained
    net's guess has a larger residual than the zero guess, so the fallback rejects it; one
    V-cycle of the rhs as the guess costs no iterations; an exact per-block mean-free correction
    of that guess -- the most a per-block network could add to it -- saves at most one more."""
    dev = WS.make_device(max_blocks=12000, windows=4096)
    try:
        dev.cnn_set_params(np.load(WS.DEFAULT_PARAMS)["params"])
        for seed in (1_000_000, 1_000_001, 1_000_002):
            pb = WS.Problem(dev, WS.sample(np.random.default_rng(seed)))
            r = WS.study_problem(dev, pb)
            print(f"[gpu] guess study seed {seed}: " +
                  ", ".join(f"{k} {v['iters']} ({v['r0']:.3g})" for k, v in r.items()), flush=True)
            assert all(v["status"] == 0 for v in r.values())
            assert r["net_fallback"]["iters"] <= r["zero"]["iters"]
            assert r["net"]["r0"] > 1.0   # the trained net's guess is worse than zero: rejected
            assert r["x1"]["iters"] <= r["zero"]["iters"]
            assert r["x1+block_detail"]["iters"] >= r["x1"]["iters"] - 1
    finally:
        dev.close()
