"""Pins of the oracle's exact line-conditional moves (CPU only).

* `logseg` against quadrature of ∫₀ᴸ e^{−gs} ds (atol 1e-4) and the mean of `trexp_offset`
  against the quadrature mean (within 0.01·L), at the reference's (g, L) pairs
  (test/test_local.jl:43-56);
* the recorded-term ΔC(t) of every family equals the change of the oracle's global energy after
  the move is committed (test/test_local.jl:120: atol 1e-12·|op| + 1e-9·|Δ|), and the caches
  stay equal to a rebuild (:122-125);
* the exact line sampler: 40000 draws of t for a fixed move against the CDF of exp(−β ΔC(t)) on
  a 20001-point quadrature grid, KS < 0.015 at the 1..99 % quantiles (:126-171), for every
  family (the storage families are two-step block moves with storage secondaries);
* whole sweeps: the caches equal a rebuild, no improper lines and no point-cap cuts.
"""
import numpy as np
import pytest

import oracle
from conftest import random_state
from esmsampling import synth

FAMS = {"gen": 0, "theta": 1, "soc": 2, "spair": 3}


@pytest.mark.parametrize("g,L", [(3.0, 0.7), (-2.0, 1.3), (1e-12, 2.0), (-40.0, 0.5)])
def test_segment_helpers(g, L):
    oracle.load()
    lib = oracle.load()
    xs = np.linspace(0, L, 200001)
    num = np.log(np.sum(np.exp(-g * xs)) * (xs[1] - xs[0]))
    assert abs(lib.orc_logseg(L, g * L) - num) < 1e-4
    u = (np.arange(200000) + 0.5) / 200000
    s = np.array([lib.orc_trexp_offset(g, L, x) for x in u])
    w = np.exp(-g * xs)
    mu = np.sum(xs * w) / np.sum(w)
    assert np.all((s >= 0) & (s <= L))
    assert abs(s.mean() - mu) < 0.01 * L


def _state(seed=3):
    s = synth.make_system(N=8, L=12, T=24, dt=1.0, seed=seed)
    R = 2
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    v[0] = rng.normal(0, 400.0, (s.N, s.T, R))
    v[1] = np.clip(v[1] + rng.normal(0, 0.05, v[1].shape), -0.02, 1.02)
    v[2] = rng.uniform(0.0, 1.0, v[2].shape)
    v[3] = rng.uniform(0.0, 1.0, v[3].shape)
    k = s.k0.copy()
    return s, v, k, x


@pytest.mark.parametrize("fam", list(FAMS))
def test_incremental_cost_equals_global(fam):
    s, v, k, x = _state()
    o = oracle.Oracle(s)
    parts, _, r, F = o.energy(v, k, x)
    rng = np.random.default_rng(4)
    for trial in range(12):
        j = int(rng.integers(2))
        n = int(rng.integers(s.N))
        tau = int(rng.integers(s.T))
        st = int(rng.integers(2))
        t = float(rng.normal(0, 0.3 if fam != "theta" else 300.0))
        dc = o.move_cost(v, k, r, F, j, FAMS[fam], st, n, tau, [t])[0]
        v2, r2, F2 = v.copy(), r.copy(), F.copy()
        o.move_apply(v2, k, r2, F2, j, FAMS[fam], st, n, tau, t)
        p2, _, r_, F_ = o.energy(v2, k, x)
        op0, op1 = parts[j, 1], p2[j, 1]
        assert abs((op1 - op0) - dc) <= 1e-12 * abs(op0) + 1e-9 * abs(dc)
        # every other replica is untouched; the incremental caches equal a rebuild
        assert np.array_equal(p2[1 - j], parts[1 - j])
        assert np.max(np.abs(r2 - r_)) < 1e-6 and np.max(np.abs(F2 - F_)) < 1e-6


@pytest.mark.parametrize("fam,site", [("gen", (0, 3, 7)), ("theta", (0, 2, 7)), ("soc", (1, 4, 7)),
                                      ("spair", (0, 5, 7)), ("spair", (1, 1, 7))])
def test_line_draws_against_quadrature(fam, site):
    s, v, k, x = _state(seed=5)
    o = oracle.Oracle(s)
    _, _, r, F = o.energy(v, k, x)
    st, n, tau = site
    j = 0
    rng = np.random.default_rng(9)
    U = rng.random((40000, 2))
    ts = np.empty(len(U))
    for q, (u1, u2) in enumerate(U):
        t, status = o.move_draw(v, k, r, F, j, FAMS[fam], st, n, tau, u1, u2)
        assert status == 0
        ts[q] = t
    lo, hi = np.quantile(ts, 0.0005), np.quantile(ts, 0.9995)
    grid = np.linspace(lo - (hi - lo), hi + (hi - lo), 20001)
    lp = -s.beta_v * o.move_cost(v, k, r, F, j, FAMS[fam], st, n, tau, grid)
    w = np.exp(lp - lp.max())
    cdf = np.cumsum(w) / w.sum()
    tss = np.sort(ts)
    qs = np.quantile(ts, np.arange(1, 100) / 100)
    femp = np.searchsorted(tss, qs, side="right") / len(ts)
    fq = cdf[np.clip(np.searchsorted(grid, qs, side="right") - 1, 0, len(grid) - 1)]
    assert np.max(np.abs(femp - fq)) < 0.015


def test_sweeps_keep_caches_and_stats():
    s = synth.make_system(N=10, L=16, T=48, dt=1.0, seed=8)
    R = 3
    v, k, x = random_state(s, R, seed=12, spread=False)
    o = oracle.Oracle(s)
    v2, r2, F2, stats = o.sweeps(v, k, 2, key=(3, 4), counter=10, x=x, threads=2, p_xpair=0.1, xlong_moves=8)
    _, _, r_, F_ = o.energy(v2, k, x)
    assert np.max(np.abs(r2 - r_)) < 1e-6 * np.abs(s.load).max()
    assert np.max(np.abs(F2 - F_)) < 1e-6 * np.abs(F_).max()
    col, nc = o.colouring()
    T, N = s.T, s.N
    ntheta = sum(1 for n in range(N))
    assert stats[0, 0] == 2 * N * T * R                 # every gas atom
    assert stats[1, 0] == 2 * ntheta * T * R             # every potential atom
    assert stats[2, 0] == 2 * 2 * N * T * R              # both storage carriers
    assert np.all(stats[:, 2] == 0) and np.all(stats[:, 3] == 0)
    assert np.all(stats[:, 1] > 0.5 * stats[:, 0])
    # the colouring is a distance-2 colouring
    nbr = [set() for _ in range(N)]
    for a, b in zip(s.line_bus0, s.line_bus1):
        nbr[a].add(int(b))
        nbr[b].add(int(a))
    for n in range(N):
        for m in nbr[n] | set().union(*[nbr[a] for a in nbr[n]]):
            if m != n:
                assert col[m] != col[n]


def test_sweeps_independent_of_thread_count():
    s = synth.make_system(N=8, L=12, T=24, dt=1.0, seed=9)
    R = 4
    v, k, x = random_state(s, R, seed=2, spread=False)
    o = oracle.Oracle(s)
    a = o.sweeps(v, k, 1, key=(1, 1), x=x, threads=1)
    b = o.sweeps(v, k, 1, key=(1, 1), x=x, threads=4)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[3], b[3])
