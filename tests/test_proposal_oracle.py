"""The proposal network's CPU oracle, pinned without the GPU.

* BF16 rounding against hand-worked values (ties to even).
* The analytic gradient (float64, no rounding) against central finite
  differences for parameters of every layer; a planted mistake in the oracle
  (a transposed conv2 weight in the backward pass) must fail the same check.
* Pooling against a direct loop.
"""

import numpy as np
import pytest

from oracle import proposal_oracle as P


def test_bf16_round_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7 + 2 ** -8, -2.5,
                  3.0e38, 1e-30], np.float32)
    got = P.bf16_round(x)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -6, -2.5, 2.9975729e38,
                     1.0005271e-30], np.float32)
    np.testing.assert_allclose(got[:5], want[:5], rtol=0, atol=0)
    # Only the top 16 bits survive.
    assert np.all(got.view(np.uint32) & 0xFFFF == 0)


def _tiles(rng, B):
    z = rng.normal(0.0, 1.0, size=(B, 16, 1024))
    # A bright drifting track so that deeper ReLUs are active near the max.
    for b in range(B):
        c0 = int(rng.integers(100, 900))
        for t in range(16):
            z[b, t, c0 + 2 * t] += 6.0
    return z


def _params(rng):
    p = np.zeros(P.NUM_PARAMS)
    p[P.OFF_W1:P.OFF_B1] = rng.uniform(-1, 1, P.C1 * 15) * np.sqrt(6 / 15)
    p[P.OFF_W2:P.OFF_B2] = rng.uniform(-1, 1, P.C2 * 15 * P.C1) * np.sqrt(6 / 240)
    p[P.OFF_W3:P.OFF_B3] = rng.uniform(-1, 1, P.C3 * 15 * P.C2) * np.sqrt(6 / 480)
    p[P.OFF_WH:P.OFF_BH] = rng.uniform(-1, 1, P.C3) * np.sqrt(6 / 32)
    p[P.OFF_B1:P.OFF_W2] = rng.uniform(0, 0.1, P.C1)
    p[P.OFF_B2:P.OFF_W3] = rng.uniform(0, 0.1, P.C2)
    p[P.OFF_B3:P.OFF_WH] = rng.uniform(0, 0.1, P.C3)
    return p


def _fd_check(grad_fn, seed=3):
    rng = np.random.default_rng(seed)
    z = _tiles(rng, 2)
    y = np.array([1.0, 0.0])
    p = _params(rng)
    _, g = grad_fn(z, y, p, 3.0)
    # Parameters with the largest gradients in each layer, plus biases and head.
    picks = []
    for lo, hi in [(P.OFF_W1, P.OFF_B1), (P.OFF_B1, P.OFF_W2), (P.OFF_W2, P.OFF_B2),
                   (P.OFF_B2, P.OFF_W3), (P.OFF_W3, P.OFF_B3), (P.OFF_B3, P.OFF_WH),
                   (P.OFF_WH, P.OFF_BH), (P.OFF_BH, P.NUM_PARAMS)]:
        seg = np.abs(g[lo:hi])
        picks += list(lo + np.argsort(seg)[-3:])
    worst = 0.0
    for i in picks:
        eps = 1e-5
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        lp = P.loss_and_dz(P.forward(z, pp, bf16=False)[0], y, 3.0)[0]
        lm = P.loss_and_dz(P.forward(z, pm, bf16=False)[0], y, 3.0)[0]
        num = (lp - lm) / (2 * eps)
        worst = max(worst, abs(num - g[i]) / max(1e-8, abs(num), abs(g[i])))
    return worst


def test_gradient_matches_finite_differences():
    assert _fd_check(lambda z, y, p, w: P.gradient(z, y, p, w, bf16=False)) < 1e-4


def test_planted_mistake_fails_finite_differences(monkeypatch):
    # The backward pass alone gets conv2's taps reversed (gradient() unpacks
    # the weights for its backward pass first, then forward() unpacks them
    # again); finite differences of the true loss must expose it.
    real_unpack = P.unpack
    calls = []

    def planted_unpack(params, bf16=True):
        w1, b1, w2, b2, w3, b3, wh, bh = real_unpack(params, bf16)
        calls.append(1)
        if len(calls) % 2 == 1:
            w2 = w2[:, ::-1, :]
        return w1, b1, w2, b2, w3, b3, wh, bh

    def planted_gradient(z, y, p, w):
        monkeypatch.setattr(P, "unpack", planted_unpack)
        try:
            return P.gradient(z, y, p, w, bf16=False)
        finally:
            monkeypatch.setattr(P, "unpack", real_unpack)

    assert _fd_check(planted_gradient) > 1e-2


def test_pool_matches_direct_loop():
    rng = np.random.default_rng(7)
    data = rng.exponential(1.0, size=(16, 8192)).astype(np.float32)
    data[3, 100] = 5000.0
    med, sd = 15.7, 4.2
    got = P.pool(data, med, sd)
    assert got.shape == (2, 16, 1024)
    for blk, t, j in [(0, 0, 0), (0, 3, 25), (1, 15, 1023), (1, 7, 512)]:
        c = blk * 4096 + 4 * j
        s = np.float32((data[t, c] + data[t, c + 1]) + (data[t, c + 2] + data[t, c + 3]))
        v = np.float32(2.0) * (s - np.float32(0.25) * np.float32(med)) / np.float32(sd)
        v = min(max(v, -4.0), 32.0)
        assert got[blk, t, j] == P.bf16_round(np.array([v], np.float32))[0]
    assert got[0, 3, 25] == 32.0  # clamped
