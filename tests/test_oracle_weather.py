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
