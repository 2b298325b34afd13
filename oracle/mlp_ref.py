"""CPU oracle of the learned step scheduler's network (test infrastructure only).

The network (inputs -> 128 -> 128 -> 128 -> 16, GELU in tanh form) and its
loss have no counterpart in KPP; this module restates the device kernels of
kpp_amd/csrc/mlp.hip in float64 numpy so tests can check the device's
forward pass and gradients.

Rounding model: the device stores weights, inputs, hidden activations and the
back-propagated dL/dZ (and dL/dO) in BF16 and accumulates every product in
FP32.  With ``emulate=True`` this oracle rounds the same stored quantities to
BF16 (round to nearest even) and computes everything else in float64, so a
device/oracle difference is FP32-vs-FP64 accumulation only.  With
``emulate=False`` it is the plain float64 network, whose gradient is pinned
against central finite differences of its own loss (tests/test_mlp_oracle.py).
"""
from __future__ import annotations

import numpy as np

HID = 128
NOUT = 16
NBUCKET = 8
GELU_C = 0.7978845608028654


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layout(in_pad):
    """Offsets (w_off[4], b_off[4], total) of the flat parameter vector."""
    kin = [in_pad, HID, HID, HID]
    nout = [HID, HID, HID, NOUT]
    w_off, b_off, o = [], [], 0
    for l in range(4):
        w_off.append(o)
        o += nout[l] * kin[l]
        b_off.append(o)
        o += nout[l]
    return w_off, b_off, o


def unpack(flat, in_pad):
    kin = [in_pad, HID, HID, HID]
    nout = [HID, HID, HID, NOUT]
    w_off, b_off, _ = layout(in_pad)
    flat = np.asarray(flat, np.float64)
    return [(flat[w_off[l]:w_off[l] + nout[l] * kin[l]].reshape(nout[l], kin[l]),
             flat[b_off[l]:b_off[l] + nout[l]]) for l in range(4)]


def pack(layers, in_pad):
    w_off, b_off, total = layout(in_pad)
    flat = np.zeros(total)
    for l, (W, b) in enumerate(layers):
        flat[w_off[l]:w_off[l] + W.size] = W.ravel()
        flat[b_off[l]:b_off[l] + b.size] = b
    return flat


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(GELU_C * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(GELU_C * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_C * (1.0 + 3.0 * 0.044715 * z * z)


def features(C, temp, sun, mu, sd, cfloor, nspec, in_pad):
    """Inputs of each cell: standardised log10 concentrations, TEMP, SUN, zero padding."""
    n = C.shape[1]
    raw = np.zeros((n, nspec + 2))
    raw[:, :nspec] = np.log10(np.maximum(C[:nspec].T, cfloor)).astype(np.float32)
    raw[:, nspec] = np.asarray(temp, np.float32)
    raw[:, nspec + 1] = 0.0 if sun is None else np.asarray(sun, np.float32)
    X = np.zeros((n, in_pad))
    X[:, :nspec + 2] = bf16((raw.astype(np.float32) - mu[:nspec + 2]) / sd[:nspec + 2])
    return X


def forward(flat, X, in_pad, emulate=True):
    r = bf16 if emulate else (lambda a: a)
    layers = unpack(flat, in_pad)
    H = r(X)
    Zs, Hs = [], [H]
    for l in range(3):
        W, b = layers[l]
        Z = H @ r(W).T + b
        H = r(gelu(Z))
        Zs.append(Z)
        Hs.append(H)
    W, b = layers[3]
    O = H @ r(W).T + b
    return O, Zs, Hs


def loss(O, y, bucket, wreg):
    """Summed loss: w*(o0-y)^2 + cross-entropy of the bucket logits."""
    lg = O[:, 1:1 + NBUCKET]
    mx = lg.max(1, keepdims=True)
    lse = np.log(np.exp(lg - mx).sum(1)) + mx[:, 0]
    ce = lse - lg[np.arange(len(O)), bucket]
    return float((wreg * (O[:, 0] - y) ** 2 + ce).sum())


def grads(flat, X, y, bucket, wreg, in_pad, emulate=True):
    """(summed loss, gradient of the mean loss in the flat parameter layout)."""
    r = bf16 if emulate else (lambda a: a)
    layers = unpack(flat, in_pad)
    O, Zs, Hs = forward(flat, X, in_pad, emulate)
    n = len(O)
    lg = O[:, 1:1 + NBUCKET]
    p = np.exp(lg - lg.max(1, keepdims=True))
    p /= p.sum(1, keepdims=True)
    dO = np.zeros_like(O)
    dO[:, 0] = 2.0 * wreg * (O[:, 0] - y) / n
    dO[:, 1:1 + NBUCKET] = p / n
    dO[np.arange(n), 1 + bucket] -= 1.0 / n
    g = [None] * 4
    d = dO
    for l in (3, 2, 1, 0):
        W, _ = layers[l]
        db = d.sum(0)
        dW = r(d).T @ Hs[l]
        g[l] = (dW, db)
        if l > 0:
            dH = r(d) @ r(W)
            d = dH * gelu_grad(Zs[l - 1])
    return loss(O, y, bucket, wreg), pack(g, in_pad)


def adam(w, g, m1, m2, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    c1 = 1 - b1 ** step
    c2 = 1 - b2 ** step
    return w - lr * (m1 / c1) / (np.sqrt(m2 / c2) + eps), m1, m2
