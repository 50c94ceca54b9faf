"""CPU pins of the learned closure's network oracle (oracle/mlp_oracle.py),
the checker of the GPU forward/backward kernels:

* BF16 rounding against hand-worked values (ties to even, carries into the
  exponent);
* the forward pass against a network whose output is known in closed form
  (identity hidden layers: y = W4 relu(x) + b4), which fails on a
  transposed weight, a wrong packing offset or a missing ReLU;
* the backward pass against central finite differences of the loss (float64,
  no BF16 rounding), which fails on a dropped ReLU mask, a wrong sign or a
  transposed operand in any layer;
* the loss mask: targets < 0 and zero weights contribute neither loss nor
  gradient.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def test_bf16_rounding_known_values():
    one = 1.0
    ulp = 2.0 ** -7  # BF16 keeps 7 explicit mantissa bits
    cases = [
        (1.0, 1.0),
        (one + ulp / 2, 1.0),                       # tie -> even (down)
        (one + 3 * ulp / 2, one + 2 * ulp),          # tie -> even (up)
        (one + ulp / 2 + 2.0 ** -20, one + ulp),     # above the tie
        (-(one + 3 * ulp / 2), -(one + 2 * ulp)),
        (2.0 - ulp / 4, 2.0),                        # carry into the exponent
        (3.0e-3, 0.0030059814453125),
        (0.0, 0.0),
    ]
    for x, want in cases:
        got = float(mo.bf16(np.float32(x)))
        assert got == want, (x, got, want)


def test_packing_offsets():
    assert mo.NPARAM == 64 * 32 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 16 * 64 + 16 == 11472
    assert mo.OFF["b4"][0] == mo.NPARAM - 16


def _identity_net(W4, b4, b1=None):
    P = np.zeros(mo.NPARAM)
    def put(name, arr):
        o, s = mo.OFF[name]
        P[o:o + arr.size] = np.asarray(arr, dtype=np.float64).reshape(-1)
    W1 = np.zeros((64, 32)); W1[np.arange(32), np.arange(32)] = 1.0
    put("W1", W1)
    if b1 is not None:
        put("b1", b1)
    put("W2", np.eye(64))
    put("W3", np.eye(64))
    put("W4", W4)
    put("b4", b4)
    return P


def test_forward_closed_form():
    rng = np.random.default_rng(1)
    W4 = rng.normal(size=(16, 64))
    b4 = rng.normal(size=16)
    X = rng.normal(size=(40, 32))
    y, _ = mo.forward(_identity_net(W4, b4), X, rb=lambda v: np.asarray(v, dtype=np.float64))
    want = np.maximum(X, 0.0) @ W4[:, :32].T + b4
    np.testing.assert_allclose(y, want, rtol=1e-13, atol=1e-13)
    # a bias in the first layer shifts the ReLU threshold; units 32.. carry
    # only relu(b1)
    b1 = rng.normal(size=64)
    y, _ = mo.forward(_identity_net(W4, b4, b1), X, rb=lambda v: np.asarray(v, dtype=np.float64))
    h = np.concatenate([np.maximum(X + b1[:32], 0.0),
                        np.broadcast_to(np.maximum(b1[32:], 0.0), (40, 32))], axis=1)
    np.testing.assert_allclose(y, h @ W4.T + b4, rtol=1e-13, atol=1e-13)


def _random_problem(seed, n=48):
    rng = np.random.default_rng(seed)
    P = rng.normal(0.0, 0.3, mo.NPARAM)
    X = np.zeros((n, 32)); X[:, :21] = rng.normal(0.0, 1.0, (n, 21))
    T = np.zeros((n, 8)); T[:, :7] = rng.uniform(0.0, 1.0, (n, 7))
    T[:, :7][rng.random((n, 7)) < 0.2] = -1.0
    T[:, 7] = rng.uniform(0.5, 2.0, n) * (rng.random(n) < 0.85)
    return rng, P, X, T


def test_backward_matches_finite_differences():
    rng, P, X, T = _random_problem(2)
    scale = 1.0 / X.shape[0]
    _, G = mo.loss_and_grad(P, X, T, scale, bf16_round=False)
    h = 1e-6
    picks = []
    for name, (o, s) in mo.OFF.items():
        size = int(np.prod(s))
        picks += list(o + rng.choice(size, size=min(size, 12), replace=False))
    for k in picks:
        Pp, Pm = P.copy(), P.copy()
        Pp[k] += h
        Pm[k] -= h
        lp, _ = mo.loss_and_grad(Pp, X, T, scale, bf16_round=False)
        lm, _ = mo.loss_and_grad(Pm, X, T, scale, bf16_round=False)
        fd = (lp - lm) / (2 * h) * scale
        assert abs(fd - G[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, G[k])


def test_masked_targets_carry_no_loss_or_gradient():
    _, P, X, T = _random_problem(3, n=16)
    T2 = T.copy()
    T2[:, :7] = -1.0
    loss, G = mo.loss_and_grad(P, X, T2, 1.0, bf16_round=False)
    assert loss == 0.0 and not np.any(G)
    T3 = T.copy()
    T3[:, 7] = 0.0
    loss, G = mo.loss_and_grad(P, X, T3, 1.0, bf16_round=False)
    assert loss == 0.0 and not np.any(G)


def test_output_bias_gradient_closed_form():
    """dL/db4_o = scale * sum_s 2 w_s (y_so - t_so) over unmasked entries"""
    _, P, X, T = _random_problem(4)
    y, _ = mo.forward(P, X, rb=lambda v: np.asarray(v, dtype=np.float64))
    w = T[:, 7:8]
    m = (T[:, :7] >= 0) & (w > 0)
    want = np.sum(np.where(m, 2.0 * w * (y[:, :7] - T[:, :7]), 0.0), axis=0) * 0.25
    _, G = mo.loss_and_grad(P, X, T, 0.25, bf16_round=False)
    o = mo.OFF["b4"][0]
    np.testing.assert_allclose(G[o:o + 7], want, rtol=1e-12)
    assert not np.any(G[o + 7:o + 16])
