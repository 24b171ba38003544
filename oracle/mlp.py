"""CPU oracle of the chemistry surrogate MLP -- TEST INFRASTRUCTURE ONLY.

Network 7 -> 64 -> 64 -> 5 with tanh hidden layers and a linear output,
loss = mean over samples and outputs of (y - t)^2.  The GPU kernels feed
the matrix cores BF16 operands (weights, inputs, activations and the
backward signals) and accumulate in FP32; this restatement rounds the same
operands to BF16 (round to nearest even) and accumulates in FP64, so the
two differ by accumulation order and the tanh implementation only.

Parameter vector layout (float32, 6288 entries), rows padded to the
matrix-core tiles:
  W1 [64][16]  (inputs 7..15 zero)   b1 [64]
  W2 [64][64]                        b2 [64]
  W3 [16][64]  (outputs 5..15 zero)  b3 [16] (5..15 zero)
"""
from __future__ import annotations

import numpy as np

NIN, NH, NOUT = 7, 64, 5
KIN, KOUT = 16, 16
SIZES = [("W1", (NH, KIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (KOUT, NH)), ("b3", (KOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float32).reshape(-1) for name, _ in SIZES])


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    b = a.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _f32(a):
    return np.asarray(a).astype(np.float32).astype(np.float64)


def _id(a):
    return np.asarray(a, dtype=np.float64)


def forward(p, x, keep=False, exact=False):
    """exact=True: plain FP64 everywhere (the mathematics the rounded form
    approximates)."""
    rd, f32 = (_id, _id) if exact else (bf16, _f32)
    w = unpack(p)
    n = x.shape[0]
    X = np.zeros((n, KIN))
    X[:, :NIN] = x
    Xb = rd(X)
    z1 = Xb @ rd(w["W1"]).T + w["b1"]
    h1 = f32(np.tanh(f32(z1)))  # the kernels use 1 - 2/(exp(2z)+1), within 1e-6
    z2 = rd(h1) @ rd(w["W2"]).T + w["b2"]
    h2 = f32(np.tanh(f32(z2)))
    y = rd(h2) @ rd(w["W3"]).T + w["b3"]
    y = y[:, :NOUT]
    if keep:
        return y, (Xb, h1, h2)
    return y


def loss_grad(p, x, t, exact=False):
    """Mean squared error and its gradient with respect to the parameters."""
    rd, f32 = (_id, _id) if exact else (bf16, _f32)
    w = unpack(p)
    n = x.shape[0]
    y, (Xb, h1, h2) = forward(p, x, keep=True, exact=exact)
    r = y - t
    loss = float(np.mean(r * r))
    dY = np.zeros((n, KOUT))
    dY[:, :NOUT] = (2.0 / (n * NOUT)) * r
    dY = f32(dY)
    g = {}
    g["W3"] = rd(dY).T @ rd(h2)
    g["b3"] = dY.sum(axis=0)
    dH2 = rd(dY) @ rd(w["W3"])
    dZ2 = f32(dH2 * (1.0 - h2 * h2))
    g["W2"] = rd(dZ2).T @ rd(h1)
    g["b2"] = dZ2.sum(axis=0)
    dH1 = rd(dZ2) @ rd(w["W2"])
    dZ1 = f32(dH1 * (1.0 - h1 * h1))
    g["W1"] = rd(dZ1).T @ Xb
    g["W1"][:, NIN:] = 0.0
    g["b1"] = dZ1.sum(axis=0)
    return loss, pack(g)


def init(seed=0):
    """Glorot-uniform weights, zero biases, padded entries zero."""
    rng = np.random.default_rng(seed)
    d = {name: np.zeros(shape) for name, shape in SIZES}
    a1, a2, a3 = np.sqrt(6 / (NIN + NH)), np.sqrt(6 / (NH + NH)), np.sqrt(6 / (NH + NOUT))
    d["W1"][:, :NIN] = rng.uniform(-a1, a1, (NH, NIN))
    d["W2"][:] = rng.uniform(-a2, a2, (NH, NH))
    d["W3"][:NOUT] = rng.uniform(-a3, a3, (NOUT, NH))
    return pack(d)
