"""CPU oracle of the learned-bound network -- TEST INFRASTRUCTURE ONLY.

There is no counterpart in the reference; this numpy restatement is pinned by central finite
differences in float64 (tests/test_mlp_oracle.py). With ``bf16=True`` it reproduces the GPU's
rounding points (BF16 weights, BF16 features/activations/deltas, FP32-like accumulation in
float64), so the GPU's forward and backward kernels can be compared against it.

Network and loss (see ebgeometry_amd/csrc/mlp.hpp):
  x_hat = (p - lo) * scale - 1
  f = [x_hat, sin(w_k x_hat_c) for k < 16, c < 3],  w_k = float32(pi * 2^(0.4 k))
  h1..h4 = relu(h W + b),  y = h4 . w5 + b5  (estimates log(|d|/diag + 0.005))
  e = y - log(|d|/diag + 0.005),  loss = w e^2 (w = under_weight if e < 0 else 1),  L = mean(loss)
  prediction = (exp(y) - 0.005) * diag
Parameter order: W1 (51x128), b1, W2, b2, W3, b3, W4, b4, w5 (128), b5.
"""
from __future__ import annotations

import numpy as np

IN, HID, FREQS = 51, 128, 16
LOSS_C, LOSS_UNDER = 0.005, 8.0
FREQ = (np.pi * 2.0 ** (0.4 * np.arange(FREQS))).astype(np.float32)
SHAPES = [(IN, HID), (HID,), (HID, HID), (HID,), (HID, HID), (HID,), (HID, HID), (HID,), (HID,), (1,)]
NPARAMS = sum(int(np.prod(s)) for s in SHAPES)


def unpack(theta):
    out, o = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(np.asarray(theta[o:o + n], dtype=np.float64).reshape(s))
        o += n
    return out


def pack(parts):
    return np.concatenate([np.asarray(p, np.float64).ravel() for p in parts])


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def features(pts, lo, scale):
    p = np.asarray(pts, np.float32)
    x = (p - np.asarray(lo, np.float32)) * np.asarray(scale, np.float32) - np.float32(1.0)
    cols = [x[:, 0], x[:, 1], x[:, 2]]
    for k in range(FREQS):
        for c in range(3):
            cols.append(np.sin(FREQ[k] * x[:, c]))
    return np.stack(cols, axis=1).astype(np.float64)


def loss_and_grad(theta, pts, target, lo, scale, diag, bf16_mode=False, under_weight=LOSS_UNDER):
    """Mean loss over the batch and its gradient (parameter order above)."""
    W1, b1, W2, b2, W3, b3, W4, b4, w5, b5 = unpack(theta)
    r = bf16 if bf16_mode else (lambda a: np.asarray(a, np.float64))
    n = pts.shape[0]
    X = r(features(pts, lo, scale))
    Ws = [r(W1), r(W2), r(W3), r(W4)]
    bs = [b1, b2, b3, b4]
    hs = [X]
    for W, b in zip(Ws, bs):
        hs.append(r(np.maximum(hs[-1] @ W + b, 0.0)))
    h4 = hs[-1]
    y = h4 @ w5 + b5[0]
    lt = np.log(np.abs(np.asarray(target, np.float64)) / diag + LOSS_C)
    e = y - lt
    w = np.where(e < 0, under_weight, 1.0)
    L = float(np.mean(w * e * e))
    g = 2.0 * w * e / n
    dw5 = h4.T @ g
    db5 = np.array([g.sum()])
    D = r(np.outer(g, w5) * (h4 > 0))
    grads = [None] * 8
    for L_ in range(3, -1, -1):
        grads[2 * L_] = hs[L_].T @ D
        grads[2 * L_ + 1] = D.sum(axis=0)
        if L_ > 0:
            D = r((D @ Ws[L_].T) * (hs[L_] > 0))
    return L, pack(grads + [dw5, db5])


def predict(theta, pts, lo, scale, diag, bf16_mode=False):
    W1, b1, W2, b2, W3, b3, W4, b4, w5, b5 = unpack(theta)
    r = bf16 if bf16_mode else (lambda a: np.asarray(a, np.float64))
    h = r(features(pts, lo, scale))
    for W, b in zip([W1, W2, W3, W4], [b1, b2, b3, b4]):
        h = r(np.maximum(h @ r(W) + b, 0.0))
    return (np.exp(h @ w5 + b5[0]) - LOSS_C) * diag


def adam_train(theta, pts, target, lo, scale, diag, batch, epochs, lr=2e-3, bf16_mode=True, under_weight=LOSS_UNDER):
    """The GPU trainer's schedule restated: Adam (beta 0.9 / 0.999, eps 1e-8, bias-corrected)
    over consecutive batches (the last one partial), learning rate decayed by a cosine from lr
    to 0.05 lr over all steps. Returns (theta, mean loss of each epoch)."""
    theta = np.array(theta, np.float64)
    m = np.zeros_like(theta)
    v = np.zeros_like(theta)
    n = pts.shape[0]
    spe = (n + batch - 1) // batch
    total = float(spe * epochs)
    k = 0
    losses = []
    for _ in range(epochs):
        acc = 0.0
        for b in range(0, n, batch):
            nb = min(batch, n - b)
            L, g = loss_and_grad(theta, pts[b:b + nb], target[b:b + nb], lo, scale, diag, bf16_mode, under_weight)
            acc += L * nb
            step_lr = lr * (0.05 + 0.95 * 0.5 * (1.0 + np.cos(np.pi * k / total)))
            k += 1
            m = 0.9 * m + 0.1 * g
            v = 0.999 * v + 0.001 * g * g
            mh = m / (1.0 - 0.9 ** k)
            vh = v / (1.0 - 0.999 ** k)
            theta = theta - step_lr * mh / (np.sqrt(vh) + 1e-8)
        losses.append(acc / n)
    return theta, losses
