"""The CPU oracle of the learned guess's network (oracle/cnn_oracle.py), pinned by finite
differences of its own loss and by known BF16 roundings."""
from __future__ import annotations

import numpy as np

from oracle import cnn_oracle as K
from sayakaamr_amd import warmstart as WS


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 3.0e-3, 0.0])
    r = K.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                          # tie -> even
    assert r[2] == 1.0 + 2.0 ** -7              # above the tie -> up
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) <= 2.0 ** -8 * 3.0e-3   # half an ulp of an 8-bit significand
    assert r[5] == 0.0


def _case(seed, n=2):
    rng = np.random.default_rng(seed)
    p = WS.init_params(seed).astype(np.float64)
    pl = K.PL
    for l in range(4):   # non-zero biases so that every ReLU sees both signs
        p[l * pl + 27 * 256:(l + 1) * pl] = rng.normal(0, 0.1, 16) * (1 if l < 3 else np.eye(16)[0])
    w = np.zeros((n, K.WIN ** 3, 16))
    w[:, :, :5] = rng.normal(0, 1, (n, K.WIN ** 3, 5))
    t = rng.normal(0, 1, (n, 512))
    return p, w, t


def test_oracle_gradient_finite_differences():
    p, w, t = _case(0)
    _, g = K.loss_and_grad_exact(p, w, t)
    rng = np.random.default_rng(1)
    for _ in range(3):
        u = rng.normal(0, 1, p.size)
        for l in range(4):   # padded entries are not parameters of the net
            blk = u[l * K.PL:l * K.PL + 27 * 256].reshape(27, 16, 16)
            cin, cout = ((5, 16), (16, 16), (16, 16), (16, 1))[l]
            blk[:, cin:, :] = 0
            blk[:, :, cout:] = 0
            u[l * K.PL + 27 * 256 + cout:(l + 1) * K.PL] = 0
        u /= np.linalg.norm(u)
        h = 1e-6   # small enough that no ReLU changes sign along the segment
        fd = (K.loss_only(p + h * u, w, t) - K.loss_only(p - h * u, w, t)) / (2 * h)
        assert abs(fd - g @ u) <= 1e-5 * max(1e-3, abs(fd)), (fd, g @ u)


def test_bf16_gradient_close_to_exact():
    p, w, t = _case(2)
    l1, g1, _ = K.loss_and_grad(p, w, t)
    l2, g2 = K.loss_and_grad_exact(p, w, t)
    assert abs(l1 - l2) <= 2e-2 * l2
    assert np.linalg.norm(g1 - g2) <= 5e-2 * np.linalg.norm(g2)


def test_adam_step():
    p, g = np.ones(4), np.array([1.0, -1.0, 0.5, 0.0])
    q, m, v = K.adam(p, g, np.zeros(4), np.zeros(4), 1, lr=0.1)
    # first step of Adam moves each parameter by lr * sign(g)
    assert np.allclose(q, p - 0.1 * np.sign(g), atol=1e-6)
