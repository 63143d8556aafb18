"""CPU oracle of the rate model (24 -> 64 -> 64 -> 1 MLP, GELU) -- TEST
INFRASTRUCTURE ONLY: the checker of the GPU forward / gradient kernels, never
part of libszx_amd.so.

The model has no counterpart in the reference (the reference's fixed-rate
mode uses an analytic model, szx_float.cc:501-584), so this oracle is pinned
by mathematics instead: with BF16 rounding switched off its analytic gradient
must match central finite differences of its own loss (tests/test_oracle.py).
With rounding on it applies BF16 round-to-nearest-even at the points the GPU
kernels do: inputs X, weights W1 and W2, activations G1, and the backward
signals dZ2 and dZ1.  Accumulations are float64 here (the GPU accumulates in
FP32 in matrix-core order), so comparisons use a tolerance.

Parameter layout (FP32): W1[64][24] | b1[64] | W2[64][64] (out, in) | b2[64]
| w3[64] | b3.
"""
from __future__ import annotations

import numpy as np

N_IN, N_H = 24, 64
SIZES = [("W1", (N_H, N_IN)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)), ("w3", (N_H,)),
         ("b3", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)
GELU_C = np.sqrt(2.0 / np.pi)
GELU_A = 0.044715


def unpack(P):
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = np.asarray(P[o:o + k], np.float64).reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SIZES])


def bf16(a):
    """Round to BF16 (nearest even) and back, elementwise."""
    a32 = np.ascontiguousarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(GELU_C * (z + GELU_A * z ** 3)))


def gelu_grad(z):
    t = np.tanh(GELU_C * (z + GELU_A * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_C * (1.0 + 3.0 * GELU_A * z * z)


def _r(a, on):
    return bf16(a) if on else np.asarray(a, np.float64)


def forward(P, X, rounding=True):
    p = unpack(P)
    X = np.asarray(X, np.float64)
    xb = _r(X, rounding)
    z1 = xb @ _r(p["W1"], rounding).T + p["b1"]
    g1b = _r(gelu(z1), rounding)
    z2 = g1b @ _r(p["W2"], rounding).T + p["b2"]
    g2 = gelu(z2)
    y = g2 @ p["w3"] + p["b3"][0]
    return y, (xb, z1, g1b, z2, g2)


def loss(P, X, T, rounding=True):
    y, _ = forward(P, X, rounding)
    return float(np.mean((y - np.asarray(T, np.float64)) ** 2))


def grad(P, X, T, rounding=True):
    """(gradient as a flat array, loss) of mean((y - T)^2)."""
    p = unpack(P)
    y, (xb, z1, g1b, z2, g2) = forward(P, X, rounding)
    n = len(y)
    e = y - np.asarray(T, np.float64)
    dy = 2.0 * e / n
    g = {"b3": np.array([dy.sum()]), "w3": g2.T @ dy}
    dz2 = dy[:, None] * p["w3"][None, :] * gelu_grad(z2)
    g["b2"] = dz2.sum(0)
    dz2b = _r(dz2, rounding)
    g["W2"] = dz2b.T @ g1b
    dg1 = dz2b @ _r(p["W2"], rounding)
    dz1 = dg1 * gelu_grad(z1)
    g["b1"] = dz1.sum(0)
    g["W1"] = _r(dz1, rounding).T @ xb
    return pack(g), float(np.mean(e * e))


def init_params(seed=0):
    """He-style initial parameters (float32)."""
    r = np.random.default_rng(seed)
    d = {"W1": r.standard_normal((N_H, N_IN)) * np.sqrt(2.0 / N_IN), "b1": np.zeros(N_H),
         "W2": r.standard_normal((N_H, N_H)) * np.sqrt(2.0 / N_H), "b2": np.zeros(N_H),
         "w3": r.standard_normal(N_H) * np.sqrt(1.0 / N_H), "b3": np.zeros(1)}
    return pack(d).astype(np.float32)
