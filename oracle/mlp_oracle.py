"""CPU oracle of the learned closure's network (numpy).

TEST INFRASTRUCTURE ONLY: the checker for the GPU forward/backward
kernels (sac_amd/csrc/sac_mlp.h, sac_closure.hip).  Only tests/ load it.

The network and the loss follow the GPU definition: 32 inputs -> 64 -> 64 ->
64 (ReLU) -> 16 outputs (linear); parameters packed W1[64][32] b1 W2[64][64]
b2 W3[64][64] b3 W4[16][64] b4; loss sum over samples and the first 7
outputs of w*(y - t)^2 where t >= 0 and w > 0.  With bf16=True every value
the GPU stores in BF16 (inputs, weights, hidden activations, the output
error, the hidden deltas) is rounded to BF16 here at the same point; sums
are exact (float64), which the GPU's FP32 accumulation follows to ~1e-6.
With bf16=False the same formulas run in float64 throughout, which is what
the finite-difference pin of the backward pass uses.
"""
from __future__ import annotations

import numpy as np

IN, H, OUT, NOUT = 32, 64, 16, 7
OFF = {}
_o = 0
for _name, _shape in (("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)),
                      ("W3", (H, H)), ("b3", (H,)), ("W4", (OUT, H)), ("b4", (OUT,))):
    OFF[_name] = (_o, _shape)
    _o += int(np.prod(_shape))
NPARAM = _o


def bf16(x):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    return {k: P[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in OFF.items()}


def forward(P, X, rb=bf16):
    p = unpack(P)
    W = {k: rb(p[k]) for k in ("W1", "W2", "W3", "W4")}
    x = rb(X)
    a1 = rb(np.maximum(x @ W["W1"].T + p["b1"], 0.0))
    a2 = rb(np.maximum(a1 @ W["W2"].T + p["b2"], 0.0))
    a3 = rb(np.maximum(a2 @ W["W3"].T + p["b3"], 0.0))
    y = a3 @ W["W4"].T + p["b4"]
    return y, (x, a1, a2, a3, W)


def loss_and_grad(P, X, T, scale, bf16_round=True):
    rb = bf16 if bf16_round else (lambda v: np.asarray(v, dtype=np.float64))
    y, (x, a1, a2, a3, W) = forward(P, X, rb)
    T = np.asarray(T, dtype=np.float64)
    w = T[:, 7:8]
    mask = (T[:, :NOUT] >= 0.0) & (w > 0.0)
    e = np.where(mask, y[:, :NOUT] - T[:, :NOUT], 0.0)
    loss = float(np.sum(w * e * e))
    d4 = np.zeros((X.shape[0], OUT))
    d4[:, :NOUT] = rb(np.where(mask, 2.0 * w * e * scale, 0.0))
    g = {}
    g["W4"] = d4.T @ a3
    g["b4"] = d4.sum(0)
    d3 = rb((d4 @ W["W4"]) * (a3 > 0))
    g["W3"] = d3.T @ a2
    g["b3"] = d3.sum(0)
    d2 = rb((d3 @ W["W3"]) * (a2 > 0))
    g["W2"] = d2.T @ a1
    g["b2"] = d2.sum(0)
    d1 = rb((d2 @ W["W2"]) * (a1 > 0))
    g["W1"] = d1.T @ x
    g["b1"] = d1.sum(0)
    G = np.zeros(NPARAM)
    for k, (o, s) in OFF.items():
        G[o:o + int(np.prod(s))] = g[k].ravel()
    return loss, G
