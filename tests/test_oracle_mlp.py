"""Pins of the learned proposal's CPU restatement (oracle/mlp_oracle.py; CPU only).

* BF16 rounding: known answers (exact values, ties to even, the largest finite values);
* the proposal density: per hour a mixture (softmax weights, clamped log sds), summed over the
  hours, against scipy's normal density;
* the head's gradient d(mean −log q)/d out against central differences of the FP64 loss
  (weights, means inside and outside the log-sd clamp);
* the whole backward pass (exact FP64 mode) against central differences of the network's loss
  in every layer's weights and biases.
"""
import numpy as np
import pytest

import mlp_oracle as mo


def test_bf16_known_answers():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, -2.5, 0.0, 3.0e38, 1.0 + 2.0 ** -7])
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -6, 1.0, -2.5, 0.0, 1.765625 * 2.0 ** 127, 1.0 + 2.0 ** -7])
    got = mo.bf16(x)
    np.testing.assert_allclose(got, want, rtol=1e-7)
    assert mo.bf16(np.array([1.0 + 2.0 ** -8]))[0] == 1.0          # tie to even (mantissa ...0)
    assert mo.bf16(np.array([1.0 + 2.0 ** -7 + 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6   # tie to even (...1 → up)


def test_logq_against_scipy_mixtures():
    """log q(y|c) is the sum over the hours of log Σ_k w_kb N(y_b; μ_kb, σ_kb), weights the
    softmax over k per hour; scipy's normal density, hour by hour."""
    from scipy.stats import norm
    rng = np.random.default_rng(5)
    K, B, M = 3, 4, 5
    out = rng.normal(0, 0.7, (M, 64))
    out[1, 2 * K * B + 2] = -9.0                    # clamped to −7
    mu0 = rng.uniform(0, 1, (M, B))
    y = mu0 + rng.normal(0, 0.3, (M, B))
    lq, resp = mo.logq(out, mu0, y, K, B)
    for i in range(M):
        tot = 0.0
        for b in range(B):
            lg = np.array([out[i, k * B + b] for k in range(K)])
            w = np.exp(lg - lg.max())
            w /= w.sum()
            mu = np.array([mu0[i, b] + out[i, K * B + k * B + b] for k in range(K)])
            sd = np.exp(np.clip([out[i, 2 * K * B + k * B + b] for k in range(K)], -7, 2))
            dens = w * norm.pdf(y[i, b], mu, sd)
            tot += np.log(dens.sum())
            np.testing.assert_allclose(resp[i, :, b], dens / dens.sum(), rtol=1e-10)
        assert lq[i] == pytest.approx(tot, rel=1e-12)


@pytest.mark.parametrize("K,B", [(1, 3), (3, 5)])
def test_head_gradient_finite_differences(K, B):
    rng = np.random.default_rng(K * 10 + B)
    M, dout = 6, 64
    out = rng.normal(0, 0.5, (M, dout))
    out[:, 2 * K * B:3 * K * B] = rng.uniform(-3, 1.5, (M, K * B))
    out[0, 2 * K * B] = 2.5                         # one log sd beyond the clamp (zero gradient)
    mu0 = rng.uniform(0, 1, (M, B))
    y = mu0 + rng.normal(0, 0.3, (M, B))
    loss, g = mo.head_grad(out, mu0, y, K, B, dout)
    assert loss == pytest.approx(-mo.logq(out, mu0, y, K, B)[0].sum())
    f = lambda o: -mo.logq(o, mu0, y, K, B)[0].mean()
    h = 1e-6
    num = np.zeros_like(out)
    for i in range(M):
        for q in range(3 * K * B):
            e = np.zeros_like(out)
            e[i, q] = h
            num[i, q] = (f(out + e) - f(out - e)) / (2 * h)
    np.testing.assert_allclose(g, num, atol=1e-7, rtol=1e-5)
    assert g[0, 2 * K * B] == 0.0
    assert np.all(g[:, 3 * K * B:] == 0)


def test_backward_finite_differences():
    rng = np.random.default_rng(3)
    din, dh, dout, K, B, M = 10, 7, 32, 2, 5, 9          # small widths: exact FP64 mode
    dims = [(din, dh), (dh, dh), (dh, dh), (dh, dout)]
    layers = [(rng.normal(0, 0.6, (fo, fi)), rng.normal(0, 0.2, fo)) for fi, fo in dims]
    X = rng.normal(0, 1, (M, din))
    mu0 = rng.uniform(0, 1, (M, B))
    y = mu0 + rng.normal(0, 0.2, (M, B))
    out, acts = mo.forward(layers, X, rnd=mo.exact)
    _, dg = mo.head_grad(out, mu0, y, K, B, dout)
    grads = mo.backward(layers, acts, dg, rnd=mo.exact)
    h = 1e-4                 # the loss is O(10): smaller steps lose digits; no ReLU kink within 1e-3
    for li in range(4):
        W, b = layers[li]
        for (arr, gi) in ((W, grads[li][0]), (b, grads[li][1])):
            for idx in list(np.ndindex(arr.shape))[:: max(1, arr.size // 12)]:
                old = arr[idx]
                arr[idx] = old + h
                lp = mo.loss_fp64(layers, X, y, mu0, K, B)
                arr[idx] = old - h
                lm = mo.loss_fp64(layers, X, y, mu0, K, B)
                arr[idx] = old
                assert gi[idx] == pytest.approx((lp - lm) / (2 * h), rel=1e-4, abs=1e-7), (li, idx)
