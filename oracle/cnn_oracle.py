"""CPU oracle of the learned initial guess's network (test infrastructure only).

Restates, in numpy with float64 sums, the per-block CNN that the device trains and runs
(sayakaamr_amd/csrc/dev/cnn.h): four 3x3x3 'same' convolutions over a 12^3 window with 16
channels, ReLU after the first three, MSE on the central 8^3 of output channel 0.  BF16
rounding is applied where the device stores BF16 (inputs, weights, activations, activation
gradients), so that the device's FP32 sums differ from these only by summation order.

The reference has no learned component; this oracle is pinned by a finite-difference check
of its own gradients (tests/test_cnn_oracle.py) and then used to check the device's forward
output and parameter gradients.
"""
from __future__ import annotations

import numpy as np

WIN, C, NTAP, NLAYER = 12, 16, 27, 4
PL = NTAP * C * C + C            # parameters per layer (padded), the device layout
NP = NLAYER * PL
OFFS = [(t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1) for t in range(NTAP)]   # (dx, dy, dz) per tap


def bf16(x):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    """params [NP] -> per layer (W [27][ci][co], b [co])."""
    out = []
    for l in range(NLAYER):
        p = params[l * PL:(l + 1) * PL]
        out.append((p[:NTAP * C * C].reshape(NTAP, C, C), p[NTAP * C * C:]))
    return out


def shift(x, dx, dy, dz):
    """y[z][y][x] = x[z+dz][y+dy][x+dx] inside the window, 0 outside.  x: [n][z][y][x][c]."""
    y = np.zeros_like(x)
    sx = slice(max(0, -dx), WIN - max(0, dx))
    sy = slice(max(0, -dy), WIN - max(0, dy))
    sz = slice(max(0, -dz), WIN - max(0, dz))
    tx = slice(max(0, dx), WIN - max(0, -dx))
    ty = slice(max(0, dy), WIN - max(0, -dy))
    tz = slice(max(0, dz), WIN - max(0, -dz))
    y[:, sz, sy, sx, :] = x[:, tz, ty, tx, :]
    return y


def conv(x, w):
    """Y[v][co] = sum_{tap, ci} X[v + off(tap)][ci] W[tap][ci][co]."""
    out = np.zeros(x.shape[:4] + (C,))
    for t, (dx, dy, dz) in enumerate(OFFS):
        out += np.einsum("nzyxc,cd->nzyxd", shift(x, dx, dy, dz), w[t])
    return out


def conv_t(dp, w):
    """dX[v][ci] = sum_{tap, co} dP[v - off(tap)][co] W[tap][ci][co]."""
    out = np.zeros(dp.shape[:4] + (C,))
    for t, (dx, dy, dz) in enumerate(OFFS):
        out += np.einsum("nzyxd,cd->nzyxc", shift(dp, -dx, -dy, -dz), w[t])
    return out


def forward(params, windows):
    """windows [n][NV][C] (as stored) -> activations list (BF16 values) and the FP32 output."""
    layers = unpack(params)
    n = windows.shape[0]
    x = bf16(windows).reshape(n, WIN, WIN, WIN, C)
    acts = [x]
    for l, (w, b) in enumerate(layers):
        z = conv(acts[-1], bf16(w)) + b
        if l < NLAYER - 1:
            acts.append(bf16(np.maximum(z, 0.0)))
        else:
            out = z[..., 0].astype(np.float32).astype(np.float64)
    return acts, out


def loss_and_grad(params, windows, targets):
    """MSE over the central 8^3 of channel 0 and its gradient in the parameter layout."""
    layers = unpack(params)
    n = windows.shape[0]
    acts, out = forward(params, windows)
    cen = out[:, 2:10, 2:10, 2:10]
    t = targets.reshape(n, 8, 8, 8)
    e = cen - t
    loss = float(np.mean(e * e))
    dy = np.zeros(out.shape + (C,))
    dy[:, 2:10, 2:10, 2:10, 0] = 2.0 * e / (n * 512)
    dy = bf16(dy)
    grad = np.zeros(NP)
    for l in range(NLAYER - 1, -1, -1):
        w, _ = layers[l]
        if l < NLAYER - 1:
            dp = np.where(acts[l + 1] > 0, dy, 0.0)
        else:
            dp = dy
        x = acts[l]
        gw = np.zeros((NTAP, C, C))
        for tp, (dx, ddy, dz) in enumerate(OFFS):
            gw[tp] = np.einsum("nzyxc,nzyxd->cd", shift(x, dx, ddy, dz), dp)
        grad[l * PL:l * PL + NTAP * C * C] = gw.ravel()
        grad[l * PL + NTAP * C * C:(l + 1) * PL] = dp.sum(axis=(0, 1, 2, 3))
        if l > 0:
            dy = bf16(conv_t(dp, bf16(w)))
    return loss, grad, out


def loss_only(params, windows, targets):
    """The same loss in float64 without any BF16 rounding (for finite differences)."""
    layers = unpack(params)
    n = windows.shape[0]
    x = windows.reshape(n, WIN, WIN, WIN, C).astype(np.float64)
    for l, (w, b) in enumerate(layers):
        z = conv(x, w) + b
        x = np.maximum(z, 0.0) if l < NLAYER - 1 else z
    e = x[:, 2:10, 2:10, 2:10, 0] - targets.reshape(n, 8, 8, 8)
    return float(np.mean(e * e))


def loss_and_grad_exact(params, windows, targets):
    """Float64 loss and gradient without BF16 rounding (the formulas the finite-difference
    check pins)."""
    layers = unpack(params)
    n = windows.shape[0]
    x = windows.reshape(n, WIN, WIN, WIN, C).astype(np.float64)
    acts = [x]
    for l, (w, b) in enumerate(layers):
        z = conv(acts[-1], w) + b
        acts.append(np.maximum(z, 0.0) if l < NLAYER - 1 else z)
    e = acts[-1][:, 2:10, 2:10, 2:10, 0] - targets.reshape(n, 8, 8, 8)
    dy = np.zeros(acts[-1].shape)
    dy[:, 2:10, 2:10, 2:10, 0] = 2.0 * e / (n * 512)
    grad = np.zeros(NP)
    for l in range(NLAYER - 1, -1, -1):
        w, _ = layers[l]
        dp = np.where(acts[l + 1] > 0, dy, 0.0) if l < NLAYER - 1 else dy
        gw = np.zeros((NTAP, C, C))
        for tp, (dx, ddy, dz) in enumerate(OFFS):
            gw[tp] = np.einsum("nzyxc,nzyxd->cd", shift(acts[l], dx, ddy, dz), dp)
        grad[l * PL:l * PL + NTAP * C * C] = gw.ravel()
        grad[l * PL + NTAP * C * C:(l + 1) * PL] = dp.sum(axis=(0, 1, 2, 3))
        if l > 0:
            dy = conv_t(dp, w)
    return float(np.mean(e * e)), grad


def adam(p, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
