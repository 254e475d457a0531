"""CPU oracle of the pass-limit network (test infrastructure only: the GPU
kernels in bigwhoop_amd/csrc/skipnet.hip are checked against it; nothing in
the product imports this file).

The network has no counterpart in the reference; it predicts, per codeblock,
how many coding passes rate control keeps (in units of 8 passes).
16 -> 64 -> 64 -> 1 with ReLU; parameter vector (fp32):
W1[64][16] b1[64] W2[64][64] b2[64] w3[64] b3 mu[16] sc[16].

`forward`/`grad` with bf16=True restate the GPU arithmetic: inputs,
first-layer activations and the back-propagated matrices are rounded to
bfloat16 (nearest even) where the kernels feed the matrix cores, sums run
in fp32. With bf16=False everything runs in float64; that variant is pinned
against central finite differences in tests/test_skipnet_oracle.py.
"""
from __future__ import annotations

import numpy as np

NF, NH = 16, 64
W1, B1, W2, B2, W3, B3, MU, SC = 0, 1024, 1088, 5184, 5248, 5312, 5313, 5329
NPARAM = 5345
NTRAIN = 5313


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(w):
    w = np.asarray(w)
    return (w[W1:B1].reshape(NH, NF), w[B1:W2], w[W2:B2].reshape(NH, NH), w[B2:W3], w[W3:B3], w[B3],
            w[MU:SC], w[SC:SC + NF])


def init_weights(seed: int, mu=None, sc=None) -> np.ndarray:
    """He-initialised weights, zero biases, given feature normalisation."""
    rng = np.random.default_rng(seed)
    w = np.zeros(NPARAM, np.float32)
    w[W1:B1] = rng.normal(0, np.sqrt(2.0 / NF), NH * NF)
    w[W2:B2] = rng.normal(0, np.sqrt(2.0 / NH), NH * NH)
    w[W3:B3] = rng.normal(0, np.sqrt(1.0 / NH), NH)
    w[MU:SC] = 0 if mu is None else mu
    w[SC:SC + NF] = 1 if sc is None else sc
    return w


def forward(w, X, bf=True):
    """Network outputs y (units of 8 passes) and the intermediate values."""
    W1m, b1, W2m, b2, w3, b3, mu, sc = unpack(w)
    if bf:
        X = np.asarray(X, np.float32)
        Xn = bf16((X - mu.astype(np.float32)) * sc.astype(np.float32))
        Z1 = Xn @ bf16(W1m).T + b1
        H1 = bf16(np.maximum(Z1, 0))
        Z2 = H1 @ bf16(W2m).T + b2
    else:
        X = np.asarray(X, np.float64)
        Xn = (X - mu) * sc
        Z1 = Xn @ W1m.T + b1
        H1 = np.maximum(Z1, 0)
        Z2 = H1 @ W2m.T + b2
    H2 = np.maximum(Z2, 0)
    y = H2 @ w3 + b3
    return y, (Xn, Z1, H1, Z2, H2)


def loss_and_grad(w, X, t, bf=True):
    """Asymmetric square loss mean(c (y - t)^2), c = 4 below the target and
    1 above, and its gradient with respect to the 5313 trained parameters."""
    W1m, b1, W2m, b2, w3, b3, mu, sc = unpack(w)
    y, (Xn, Z1, H1, Z2, H2) = forward(w, X, bf)
    n = len(t)
    e = y - t
    c = np.where(e < 0, 4.0, 1.0)
    loss = float(np.sum(c * e * e) / n)
    dy = 2 * c * e / n
    g = np.zeros(NTRAIN, np.float64)
    g[W3:B3] = H2.T @ dy
    g[B3] = dy.sum()
    dZ2 = (Z2 > 0) * dy[:, None] * w3[None, :]
    g[B2:W3] = dZ2.sum(0)
    if bf:
        dZ2m = bf16(dZ2)
        dH1 = dZ2m @ bf16(W2m)
    else:
        dZ2m = dZ2
        dH1 = dZ2m @ W2m
    dZ1 = (Z1 > 0) * dH1
    g[B1:W2] = dZ1.sum(0)
    dZ1m = bf16(dZ1) if bf else dZ1
    g[W2:B2] = (dZ2m.T @ H1).ravel()
    g[W1:B1] = (dZ1m.T @ Xn).ravel()
    return loss, g


def limits(y, K, margin):
    """Pass limits from network outputs, as the inference kernel forms them."""
    full = np.where(K > 0, 3 * K.astype(np.int64) - 2, 1)
    p = np.clip(np.ceil(8.0 * y), 0, 255).astype(np.int64) + margin
    return np.clip(p, 1, full)
