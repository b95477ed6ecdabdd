"""CPU oracle for the surrogate network (test infrastructure only).

A numpy restatement of the GPU MLP (mcscram_amd/csrc/mlp.hip): n_in -> H x 3
(GELU) -> 1, mean squared error, with BF16 rounding at the same points as
the kernels (inputs, weights, hidden activations, backpropagated pre-
activation gradients) and FP32 elsewhere. With ``bf16=False`` every value
stays in float64 and the backward pass is checked against finite
differences of the forward pass (tests/test_surrogate.py), which pins the
gradient formulas independently of the GPU code.

The network has no counterpart in the reference; this oracle is the
mathematics of the network itself.
"""
from __future__ import annotations

import math

import numpy as np

_erf = np.vectorize(math.erf, otypes=[np.float64])


def bf16(x):
    """Round float32 values to BF16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7F800000) == 0x7F800000
    nan &= (u & 0x007FFFFF) != 0
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, ((u >> 16) | 0x40) << 16, r)
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).reshape(a.shape)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    cdf = 0.5 * (1.0 + _erf(x / math.sqrt(2.0)))
    pdf = np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)
    return cdf + x * pdf


def apply_input_map(x, scale, shift, cap):
    """The network's fixed input map, FP32: min(cap, exp(scale x + shift))
    per column (mcs_mlp_set_input_map)."""
    f = np.float32
    x = np.asarray(x, dtype=f)
    v = np.exp(np.asarray(scale, dtype=f)[None, :] * x + np.asarray(shift, dtype=f)[None, :])
    return np.minimum(np.asarray(cap, dtype=f)[None, :], v.astype(f))


def unpack(params, n_in, h):
    p = np.asarray(params)
    o = 0
    out = {}
    for name, shape in (("W1", (n_in, h)), ("b1", (h,)), ("W2", (h, h)), ("b2", (h,)),
                        ("W3", (h, h)), ("b3", (h,)), ("w4", (h,)), ("b4", (1,))):
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    assert o == p.shape[0]
    return out


def loss_and_grad(params, x, y, n_in, h, bf16_mode=True):
    """Mean squared error over the rows of x and its gradient (flat, in the
    parameter order W1, b1, W2, b2, W3, b3, w4, b4)."""
    f = np.float32 if bf16_mode else np.float64
    rnd = bf16 if bf16_mode else (lambda a: np.asarray(a, dtype=np.float64))
    P = {k: np.asarray(v, dtype=f) for k, v in unpack(params, n_in, h).items()}
    x = rnd(np.asarray(x, dtype=f))
    y = np.asarray(y, dtype=f)
    B = x.shape[0]
    W1, W2, W3, w4 = rnd(P["W1"]), rnd(P["W2"]), rnd(P["W3"]), rnd(P["w4"])
    z1 = (x @ W1).astype(f) + P["b1"]
    a1 = rnd(gelu(z1).astype(f))
    z2 = (a1 @ W2).astype(f) + P["b2"]
    a2 = rnd(gelu(z2).astype(f))
    z3 = (a2 @ W3).astype(f) + P["b3"]
    a3 = rnd(gelu(z3).astype(f))
    yhat = (a3 @ w4).astype(f) + P["b4"][0]
    e = yhat - y
    loss = float(np.mean(e.astype(np.float64) ** 2))
    d = (2.0 * e / B).astype(f)
    g = {}
    g["w4"] = a3.T @ d
    g["b4"] = np.array([d.sum()], dtype=f)
    dz3 = rnd((d[:, None] * w4[None, :] * gelu_grad(z3)).astype(f))
    g["W3"] = a2.T @ dz3
    g["b3"] = dz3.sum(axis=0)
    dz2 = rnd(((dz3 @ W3.T).astype(f) * gelu_grad(z2)).astype(f))
    g["W2"] = a1.T @ dz2
    g["b2"] = dz2.sum(axis=0)
    dz1 = rnd(((dz2 @ W2.T).astype(f) * gelu_grad(z1)).astype(f))
    g["W1"] = x.T @ dz1
    g["b1"] = dz1.sum(axis=0)
    flat = np.concatenate([np.asarray(g[k], dtype=np.float64).ravel()
                           for k in ("W1", "b1", "W2", "b2", "W3", "b3", "w4", "b4")])
    return loss, flat


def predict(params, x, n_in, h, bf16_mode=True):
    f = np.float32 if bf16_mode else np.float64
    rnd = bf16 if bf16_mode else (lambda a: np.asarray(a, dtype=np.float64))
    P = {k: np.asarray(v, dtype=f) for k, v in unpack(params, n_in, h).items()}
    a = rnd(np.asarray(x, dtype=f))
    for W, b in (("W1", "b1"), ("W2", "b2"), ("W3", "b3")):
        a = rnd(gelu((a @ rnd(P[W])).astype(f) + P[b]).astype(f))
    return (a @ rnd(P["w4"])).astype(f) + P["b4"][0]
