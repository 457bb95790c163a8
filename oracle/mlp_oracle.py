"""CPU oracle of the learned-reluctivity MLP (TEST INFRASTRUCTURE ONLY).

No reference counterpart: MagPackage has no learned component.  This is a
numpy restatement of the network of include/mpk_mlp.h, in FP64, with the
GPU kernels' BF16 rounding points reproduced when ``bf16=True`` (inputs,
hidden activations and backpropagated pre-activation gradients rounded to
BF16, round to nearest even, where they enter a matrix product) and no
rounding when ``bf16=False`` (the pure FP64 network, whose gradient the
tests pin by central differences).

Parameter vector layout: W1 [64 x 16], b1 [64], W2 [64 x 64], b2 [64],
W3 [64 x 64], b3 [64], w4 [64], b4 [1].

The activation is the kernels' tanh: the rational minimax form of Eigen's
generic_fast_tanh_float, x P(x^2) / Q(x^2) on x clamped to
+-7.90531110763549805 (``tanh_rational``; within 2.6e-7 of tanh).
"""
from __future__ import annotations

import numpy as np

N_IN, N_H = 16, 64
SHAPES = (("W1", (N_H, N_IN)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)),
          ("W3", (N_H, N_H)), ("b3", (N_H,)), ("w4", (N_H,)), ("b4", (1,)))
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + k], dtype=np.float64).reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def bf16(x):
    """Round to BF16 (nearest even) through FP32, returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


_TANH_P = (4.89352455891786e-03, 6.37261928875436e-04, 1.48572235717979e-05,
           5.12229709037114e-08, -8.60467152213735e-11, 2.00018790482477e-13,
           -2.76076847742355e-16)
_TANH_Q = (4.89352518554385e-03, 2.26843463243900e-03, 1.18534705686654e-04,
           1.19825839466702e-06)
_TANH_CLAMP = 7.90531110763549805


def tanh_rational(x):
    """x P(x^2) / Q(x^2) on the clamped x, in FP64 (Horner in x^2)."""
    x = np.clip(np.asarray(x, dtype=np.float64), -_TANH_CLAMP, _TANH_CLAMP)
    x2 = x * x
    p = np.zeros_like(x2)
    for c in _TANH_P[::-1]:
        p = p * x2 + c
    q = np.zeros_like(x2)
    for c in _TANH_Q[::-1]:
        q = q * x2 + c
    return x * p / q


def init_params(seed=0):
    """Glorot-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SHAPES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / (shape[0] + shape[1]))
            d[name] = rng.uniform(-lim, lim, shape)
        elif name == "w4":
            lim = np.sqrt(6.0 / (N_H + 1))
            d[name] = rng.uniform(-lim, lim, shape)
        else:
            d[name] = np.zeros(shape)
    return pack(d)


def forward(p, X, bf=True):
    q = unpack(p)
    rd = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    W = {k: rd(q[k]) for k in ("W1", "W2", "W3")}
    x = rd(X)
    act = tanh_rational if bf else np.tanh
    h1 = rd(act(x @ W["W1"].T + q["b1"]))
    h2 = rd(act(h1 @ W["W2"].T + q["b2"]))
    h3 = rd(act(h2 @ W["W3"].T + q["b3"]))
    y = h3 @ q["w4"] + q["b4"][0]
    return y, (x, h1, h2, h3, W)


def loss_grad(p, X, t, bf=True):
    """MSE loss sum (y - t)^2 / n and its gradient (NPARAM,)."""
    q = unpack(p)
    rd = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    n = X.shape[0]
    y, (x, h1, h2, h3, W) = forward(p, X, bf)
    d = y - t
    loss = float(d @ d / n)
    dy = 2.0 * d / n
    g = {"w4": h3.T @ dy, "b4": np.array([dy.sum()])}
    dz3 = dy[:, None] * q["w4"][None, :] * (1.0 - h3 * h3)
    g["b3"] = dz3.sum(axis=0)
    dz3b = rd(dz3)
    g["W3"] = dz3b.T @ h2
    dz2 = (dz3b @ W["W3"]) * (1.0 - h2 * h2)
    g["b2"] = dz2.sum(axis=0)
    dz2b = rd(dz2)
    g["W2"] = dz2b.T @ h1
    dz1 = (dz2b @ W["W2"]) * (1.0 - h1 * h1)
    g["b1"] = dz1.sum(axis=0)
    dz1b = rd(dz1)
    g["W1"] = dz1b.T @ x
    return loss, pack(g)
