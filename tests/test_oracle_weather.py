"""Pins of the oracle's weather process (CPU only).

* categorical inversion: the reference's known answers (test/test_replicas.jl:26-27);
* the regime window redraw: the law of the redrawn regimes, measured from the explicit-uniform
  map by bisection, against enumeration of the Markov-chain conditional (the reference checks
  the window conditional against enumeration, test/test_replicas.jl:139-160);
* prior draws: closed forms of the synthetic process (stationary variance σ²/(1 − φ²) of the
  anomalies with zero intercepts, mean regime sojourn 1/0.02 = 50 steps, standard-normal
  innovations);
* pCN steps keep the prior's mean and variance and never touch the regimes;
* the compensated re-dispatch: where the gas unit has the headroom (or output), the node
  residual after the weather step equals the one before; elsewhere the gas unit is at its bound.
"""
import copy
import itertools
import math

import numpy as np
import pytest

import oracle
from esmsampling import synth


def test_categorical_known_answers():
    p = [0.2, 0.0, 0.5, 0.3]
    got = [oracle.categorical(p, u) for u in (0.0, 0.1, 0.2, 0.2000001, 0.7, 0.70001, 1.0)]
    assert got == [0, 0, 0, 2, 2, 3, 3]


def _interval_lengths(f, K):
    """Lengths of the u-sets of [0, 1] on which the monotone map f takes each of K values."""
    out = np.zeros(K)
    lo, flo = 0.0, f(0.0)
    while True:
        fhi = f(1.0)
        if flo == fhi:
            out[flo] += 1.0 - lo
            return out
        a, b = lo, 1.0                  # f(a) = flo, f(b) != flo
        for _ in range(60):
            mid = 0.5 * (a + b)
            if f(mid) == flo:
                a = mid
            else:
                b = mid
        out[flo] += b - lo
        lo, flo = b, f(b)


@pytest.mark.parametrize("a,b,T", [(3, 4, 8), (1, 2, 6), (7, 8, 8), (2, 4, 7)])
def test_regime_window_against_enumeration(a, b, T):
    s = synth.make_system(N=3, L=3, T=T, seed=4)
    s = copy.copy(s)
    s.Ptr = np.array([[0.9, 0.1], [0.3, 0.7]])
    s.pi0 = np.array([0.4, 0.6])
    o = oracle.Oracle(s)
    rng = np.random.default_rng(1)
    path = rng.integers(0, 2, T).astype(np.int32)
    n = b - a + 1
    # enumeration of the conditional law of m[a..b] (1-based) given the rest of the path
    enum = {}
    for cfg in itertools.product(range(2), repeat=n):
        mm = path.copy()
        mm[a - 1:b] = cfg
        p = s.pi0[mm[0]]
        for t in range(1, T):
            p *= s.Ptr[mm[t - 1], mm[t]]
        enum[cfg] = p
    tot = sum(enum.values())
    # law of the explicit-uniform map: product of the interval lengths of each draw
    law = {}

    def rec(prefix, us):
        q = len(prefix)
        if q == n:
            law[tuple(prefix)] = np.prod(us)
            return
        def f(u):
            uu = [0.5] * n
            for i, c in enumerate(prefix):
                uu[i] = mids[i][c]
            uu[q] = u
            return int(o.regime_window_path(path, a, b, uu)[a - 1 + q])
        lens = _interval_lengths(f, 2)
        for c in range(2):
            if lens[c] > 0:
                lo = sum(lens[:c])
                mids[q] = dict(mids[q]) if q < len(mids) else {}
                mids[q][c] = lo + 0.5 * lens[c]
                rec(prefix + [c], us + [lens[c]])
    mids = [dict() for _ in range(n)]
    rec([], [])
    for cfg, p in enum.items():
        assert law.get(cfg, 0.0) == pytest.approx(p / tot, abs=1e-12)


def _zero_mean_system(T, N=6):
    s = synth.make_system(N=N, L=8, T=T, seed=5)
    s = copy.copy(s)
    s.mu = np.zeros_like(s.mu)
    return s


def test_prior_draw_closed_forms():
    s = _zero_mean_system(T=20000)
    R = 4
    m, z, x = oracle.Oracle(s).weather_init(R, key=(7, 9), counter=3)
    var = s.sigma ** 2 / (1 - s.phi ** 2)
    assert z.var() == pytest.approx(var, rel=0.03)
    assert abs(z.mean()) < 0.03
    switches = (np.diff(m, axis=0) != 0).sum()
    sojourn = m.shape[0] * R / (switches + R)
    assert sojourn == pytest.approx(1 / 0.02, rel=0.1)
    # innovations of the draw are standard normal
    e = (z[:, 1:] - s.phi * z[:, :-1]) / s.sigma
    assert abs(e.mean()) < 0.01 and e.var() == pytest.approx(1.0, rel=0.02)
    # x = sigmoid(xbar(hour of day) + z)
    hod = synth.hour_of_day(s.T, s.dt)
    xb = s.xbar[:, hod]
    np.testing.assert_allclose(x.reshape(-1, s.T, R), 1 / (1 + np.exp(-(xb[:, :, None] + z))), rtol=1e-14)


def test_pcn_keeps_the_prior():
    s = _zero_mean_system(T=4000)
    R = 8
    o = oracle.Oracle(s)
    m, z, x = o.weather_init(R, key=(1, 2), counter=0)
    v = np.zeros((4, s.N, s.T, R))
    for it in range(20):
        m2, z, x, v, moved = o.weather_step(m, z, x, v, s.k0, np.full(R, 0.5), np.zeros(R), compensate=False,
                                            key=(1, 2), counter=it + 1)
        assert np.array_equal(m2, m) and np.all(moved == 1)
    var = s.sigma ** 2 / (1 - s.phi ** 2)
    assert z.var() == pytest.approx(var, rel=0.04)
    assert abs(z.mean()) < 0.03


def test_compensation_restores_the_residual():
    s = synth.make_system(N=6, L=8, T=96, seed=6)
    R = 3
    o = oracle.Oracle(s)
    m, z, x = o.weather_init(R, key=(3, 3), counter=0)
    v = synth.heuristic_state(s, x)
    v[1] = np.clip(v[1], 0.05, 0.95)
    k = s.k0
    r0 = o.energy(v, k, x)[2]
    m2, z2, x2, v2, moved = o.weather_step(m, z, x, v, k, np.full(R, 0.8), np.full(R, 1.0), window=20,
                                           compensate=True, key=(3, 3), counter=1)
    assert np.all(moved == 2)
    r1 = o.energy(v2, k, x2)[2]
    kap = k[3][:, None, None]
    y0, y1 = v[1], v2[1]
    assert np.all(v2[[0, 2, 3]] == v[[0, 2, 3]])                 # only the gas utilization moves
    inside = (y1 > 0) & (y1 < 1)
    scale = np.abs(s.load).max()
    assert np.all(np.abs(r1 - r0)[inside] <= 1e-9 * scale)       # fully absorbed
    # where not absorbed the unit sits at a bound of its box
    outside = ~inside & (np.abs(r1 - r0) > 1e-9 * scale)
    assert np.all((np.abs(y1[outside]) < 1e-12) | (np.abs(y1[outside] - 1) < 1e-12))
    # the change of output equals the absorbed part of the residual change
    dp = kap * (np.clip(y1, 0, 1) - np.clip(y0, 0, 1))
    need_r = r1 + dp - r0
    np.testing.assert_allclose(dp[inside], need_r[inside], rtol=1e-9, atol=1e-6)


def _enumerate_posterior(s, z, j):
    """log p(z, m) for every regime path m (K^T of them) of replica j, from the model's definition
    with scipy's normal density: m_1 ~ π0, m_τ | m_{τ−1} ~ Ptr, z_1 ~ N(0, σ0²), z_τ ~
    N(φ z_{τ−1} + μ_{m_τ}(τ), σ²) per series, μ_k(τ) = M_k · (1, cos 2πτ/T, sin 2πτ/T)."""
    import itertools
    from scipy.stats import norm
    T, K = s.T, s.K
    nser = 3 * s.N
    zz = z[:, :, j]                              # [nser, T]
    tau = np.arange(1, T + 1)
    harm = np.stack([np.ones(T), np.cos(2 * np.pi * tau / T), np.sin(2 * np.pi * tau / T)])   # [3, T]
    mean_k = np.einsum("kic,ct->kit", s.mu, harm)                                           # [K, nser, T]
    em = np.zeros((K, T))
    em[:, 0] = norm.logpdf(zz[:, 0], 0.0, s.sigma0).sum()
    for k in range(K):
        em[k, 1:] = norm.logpdf(zz[:, 1:], s.phi * zz[:, :-1] + mean_k[k, :, 1:], s.sigma).sum(axis=0)
    paths = list(itertools.product(range(K), repeat=T))
    lj = np.array([np.log(s.pi0[p[0]]) + sum(np.log(s.Ptr[p[t - 1], p[t]]) for t in range(1, T))
                   + sum(em[p[t], t] for t in range(T)) for p in paths])
    return paths, lj


@pytest.mark.parametrize("K", [2, 3])
def test_ffbs_against_enumeration(K):
    """The centred FFBS of the regimes (src/weather.jl:160-174) on a 5-step horizon: the forward
    filter's normalizer equals log Σ_m p(z, m) over all K^5 paths, and 6000 draws of the path
    (one per counter) follow the exact posterior p(m | z) (total variation < 0.04, every path
    with posterior mass > 0.02 within 5 standard errors)."""
    s = synth.make_system(N=2, L=1, T=5, dt=1.0, seed=4, K=K)
    o = oracle.Oracle(s)
    R = 3
    m0, z, _ = o.weather_init(R, key=(3, 1))
    z = z * 1.5                                      # away from the prior draw: a spread posterior
    for j in range(R):
        paths, lj = _enumerate_posterior(s, z, j)
        mx = lj.max()
        logZ = mx + np.log(np.exp(lj - mx).sum())
        _, ll = o.weather_ffbs(m0, z, key=(3, 2), counter=0)
        assert ll[j] == pytest.approx(logZ, rel=1e-12, abs=1e-10)
    post = np.exp(lj - logZ)
    index = {p: q for q, p in enumerate(paths)}
    n = 6000
    counts = np.zeros(len(paths))
    mask = np.array([0, 0, 1], dtype=np.int32)      # replica 2 only
    for c in range(n):
        m1, _ = o.weather_ffbs(m0, z, mask=mask, key=(3, 4), counter=c)
        counts[index[tuple(m1[:, 2])]] += 1
        assert np.array_equal(m1[:, :2], m0[:, :2])
    emp = counts / n
    assert 0.5 * np.abs(emp - post).sum() < 0.04
    big = post > 0.02
    assert np.all(np.abs(emp[big] - post[big]) < 5 * np.sqrt(post[big] * (1 - post[big]) / n))


def test_ffbs_keeps_anomalies_and_planted_mistake_fails():
    """FFBS changes only the regimes. The enumeration pin has power: a forward filter that drops
    the transition term (log Ptr) is caught by the normalizer check."""
    s = synth.make_system(N=2, L=1, T=5, dt=1.0, seed=4, K=2)
    o = oracle.Oracle(s)
    m0, z, _ = o.weather_init(2, key=(3, 1))
    z = z * 1.5
    paths, lj = _enumerate_posterior(s, z, 0)
    # the same joint without the transitions (the planted mistake)
    lj_bad = lj - np.array([sum(np.log(s.Ptr[p[t - 1], p[t]]) for t in range(1, s.T)) for p in paths])
    _, ll = o.weather_ffbs(m0, z, key=(3, 2), counter=0)
    mx = lj_bad.max()
    assert abs(ll[0] - (mx + np.log(np.exp(lj_bad - mx).sum()))) > 1e-3
