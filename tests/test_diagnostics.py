"""Convergence diagnostics and ESS (esmsampling.diagnostics) on synthetic chains of known behaviour.

Known answers of test/test_convergence.jl:5-19 (bands on iid, AR(0.9), shifted and scaled
chains, random walks, warmup detection), the Geyer estimate in closed form on AR(1), the
inverse normal against scipy's ndtri, and brute-force small cases.
"""
import math

import numpy as np
import pytest
import scipy.special as sp

from esmsampling import diagnostics as D


def _ar(rng, n, phi):
    x = np.empty(n)
    x[0] = rng.standard_normal() / math.sqrt(1 - phi * phi)
    e = rng.standard_normal(n)
    for t in range(1, n):
        x[t] = phi * x[t - 1] + e[t]
    return x


def test_norm_ppf_matches_ndtri():
    p = np.concatenate([np.logspace(-300, -1, 400), np.linspace(0.01, 0.99, 999), 1 - np.logspace(-16, -1, 200),
                        [0.075, 0.425 + 0.5, 0.5 - 0.425, 0.5, 1e-310]])
    z = D.norm_ppf(p)
    ref = sp.ndtri(p)
    np.testing.assert_allclose(z, ref, rtol=2e-15, atol=1e-300)
    assert D.norm_ppf(np.array([0.0]))[0] == -np.inf and D.norm_ppf(np.array([1.0]))[0] == np.inf
    # symmetry Φ⁻¹(1 − p) = −Φ⁻¹(p) on exactly representable pairs
    q = np.array([0.25, 0.125, 0.0625, 2.0 ** -20])
    np.testing.assert_array_equal(D.norm_ppf(1 - q), -D.norm_ppf(q))


def test_split_and_ranks_brute_force():
    sc = D.split_chains([[1, 2, 3, 4, 5], [6, 7, 8, 9]])
    assert [list(c) for c in sc] == [[1, 2], [4, 5], [6, 7], [8, 9]]
    z = D.rank_normalize([[3.0, 1.0, 1.0], [2.0, 3.0]])
    ranks = np.array([4.5, 1.5, 1.5, 3.0, 4.5])
    ref = sp.ndtri((ranks - 0.375) / 5.25)
    np.testing.assert_allclose(np.concatenate(z), ref, rtol=1e-15)


def test_autocov_and_rhat_closed_forms():
    rng = np.random.default_rng(0)
    x = rng.standard_normal(37)
    y = x - x.mean()
    ref = [sum(y[i] * y[i + t] for i in range(37 - t)) / 37 for t in range(37)]
    np.testing.assert_allclose(D.autocov(x), ref, rtol=1e-12, atol=1e-15)
    # two chains: means 0 and 2, within variance 1 each: B = n·var([0, 2]) = 2n, W = 1
    a = np.array([-1.0, 0.0, 1.0])
    chains = [a, a + 2]
    n, W, B = 3, 1.0, 3 * 2.0
    assert D.rhat_basic(chains) == pytest.approx(math.sqrt(((n - 1) / n * W + B / n) / W), rel=1e-15)
    assert math.isnan(D.rhat_basic([a]))


def test_reference_bands_iid_ar():
    rng = np.random.default_rng(1)
    iid = D.convergence_summary([rng.standard_normal(1000) for _ in range(4)])
    assert iid["passed"] and iid["rhat"] < 1.01 and 3000 < iid["ess_bulk"] < 5000
    a9 = D.convergence_summary([_ar(rng, 2000, 0.9) for _ in range(4)])
    assert 250 < a9["ess_bulk"] < 700          # expected 8000·0.1/1.9 ≈ 420


def test_reference_bands_shifts_and_random_walk():
    rng = np.random.default_rng(2)
    loc = D.convergence_summary([rng.standard_normal(1000) + (0.5 if k == 0 else 0.0) for k in range(4)])
    assert loc["rhat"] > 1.01
    scl = D.convergence_summary([rng.standard_normal(1000) * (2.0 if k == 0 else 1.0) for k in range(4)])
    assert scl["rhat"] > 1.01                  # found by the folded (tail) R̂
    rw = D.convergence_summary([np.cumsum(rng.standard_normal(1000)) for _ in range(4)])
    assert not rw["passed"]


def test_warmup_detection_band():
    rng = np.random.default_rng(3)
    tr = [[np.concatenate([np.linspace(3 * k, 0, 100) + 0.1 * rng.standard_normal(100), rng.standard_normal(300)])
           for k in range(1, 5)]]
    w = D.warmup_end(tr, every=10)
    assert 50 <= w <= 160


def test_geyer_ess_ar1_closed_form():
    """ESS of AR(φ) = n (1 − φ)/(1 + φ) for long chains."""
    rng = np.random.default_rng(4)
    for phi in (0.0, 0.5, 0.8):
        x = _ar(rng, 100_000, phi)
        e = D.ess(x)
        assert e == pytest.approx(len(x) * (1 - phi) / (1 + phi), rel=0.06), phi
        em = D.ess_multi([x[:50_000], x[50_000:]])
        assert em == pytest.approx(len(x) * (1 - phi) / (1 + phi), rel=0.06), phi


def test_constant_quantity():
    s = D.convergence_summary([np.ones(10), np.ones(10)])
    assert s["constant"] and s["passed"] and s["ess_bulk"] == 20.0
