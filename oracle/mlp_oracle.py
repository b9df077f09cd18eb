"""TEST INFRASTRUCTURE: CPU oracle of the learned CG warm start (numpy, FP64).

Only the tests, the smoke check and the bench's CPU leg use this module; the
product path is the HIP library.  The network and its features have no
counterpart in the reference (its PCG starts from zero, pcg.h:109-130), so
they are pinned here by mathematics instead of reference vectors:
- forward / backward of the 15 -> 128 -> 128 -> 128 -> 3 GELU network, checked
  by central finite differences of the loss (tests/test_oracle_mlp.py);
- GELU through the normal CDF (scipy.special.erf), checked against its
  closed-form identities;
- the feature rows, restated from their definition in include/sdx_fem.h and
  computed from the oracle's own assembled system.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

IN, ROW, HID, OUT = 15, 16, 128, 3
O_W1 = 0
O_B1 = O_W1 + HID * IN
O_W2 = O_B1 + HID
O_B2 = O_W2 + HID * HID
O_W3 = O_B2 + HID
O_B3 = O_W3 + HID * HID
O_W4 = O_B3 + HID
O_B4 = O_W4 + OUT * HID
N_PARAMS = O_B4 + OUT


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[O_W1:O_B1].reshape(HID, IN), p[O_B1:O_W2], p[O_W2:O_B2].reshape(HID, HID), p[O_B2:O_W3],
            p[O_W3:O_B3].reshape(HID, HID), p[O_B3:O_W4], p[O_W4:O_B4].reshape(OUT, HID), p[O_B4:N_PARAMS])


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(p, X, cache=False):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    x = np.asarray(X, dtype=np.float64)[:, :IN]
    z1 = x @ W1.T + b1
    h1 = gelu(z1)
    z2 = h1 @ W2.T + b2
    h2 = gelu(z2)
    z3 = h2 @ W3.T + b3
    h3 = gelu(z3)
    y = h3 @ W4.T + b4
    if cache:
        return y, (x, z1, h1, z2, h2, z3, h3)
    return y


def loss_grad(p, X, T):
    """loss = mean over rows of |y - t|^2 and its gradient in the parameter layout."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    y, (x, z1, h1, z2, h2, z3, h3) = forward(p, X, cache=True)
    M = x.shape[0]
    d = y - np.asarray(T, dtype=np.float64)
    loss = float(np.sum(d * d) / M)
    dy = 2.0 * d / M
    g = np.zeros(N_PARAMS)
    g[O_W4:O_B4] = (dy.T @ h3).ravel()
    g[O_B4:] = dy.sum(axis=0)
    dz3 = (dy @ W4) * gelu_grad(z3)
    g[O_W3:O_B3] = (dz3.T @ h2).ravel()
    g[O_B3:O_W4] = dz3.sum(axis=0)
    dz2 = (dz3 @ W3) * gelu_grad(z2)
    g[O_W2:O_B2] = (dz2.T @ h1).ravel()
    g[O_B2:O_W3] = dz2.sum(axis=0)
    dz1 = (dz2 @ W2) * gelu_grad(z1)
    g[O_W1:O_B1] = (dz1.T @ x).ravel()
    g[O_B1:O_W2] = dz1.sum(axis=0)
    return loss, g


def features(res, ptr, col, blocks, dprev):
    """Feature rows [n][16] and the scale s of one environment, from its
    residual b (3n), BSR pattern/blocks ([nnzb][3][3], Dirichlet applied)
    and previous Newton step dprev (3n)."""
    n = len(ptr) - 1
    b = np.asarray(res, dtype=np.float64).reshape(n, 3)
    dp = np.asarray(dprev, dtype=np.float64).reshape(n, 3)
    D = np.zeros((n, 3, 3))
    for i in range(n):
        for k in range(ptr[i], ptr[i + 1]):
            if col[k] == i:
                D[i] = blocks[k]
    d = np.trace(D, axis1=1, axis2=2) / 3.0
    inv = np.where(d > 0, 1.0 / np.where(d > 0, d, 1.0), 0.0)
    y = b * inv[:, None]
    s = np.sqrt(max(np.sum(dp * dp), np.sum(y * y)) / n)
    s = s if s > 0 else 1.0
    F = np.zeros((n, ROW))
    F[:, 0:3] = y / s
    F[:, 3] = D[:, 0, 0] * inv
    F[:, 4] = D[:, 1, 1] * inv
    F[:, 5] = D[:, 2, 2] * inv
    F[:, 6] = D[:, 0, 1] * inv
    F[:, 7] = D[:, 0, 2] * inv
    F[:, 8] = D[:, 1, 2] * inv
    F[:, 9:12] = dp / s
    for i in range(n):
        nb = [col[k] for k in range(ptr[i], ptr[i + 1]) if col[k] != i]
        if nb:
            F[i, 12:15] = y[nb].mean(axis=0) / s
    return F, s
