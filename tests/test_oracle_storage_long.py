"""Pins of the oracle's long-range storage moves (CPU only).

* a plateau (a, L) moves the state of charge of its unit by exactly t at the cyclic instants
  a+1 … a+L and nothing else of v except the gas at the end steps a and a+L, which follows
  ±κs t while it is inside its box and the requirement stays positive; the residuals of the
  interior steps do not change while their instants stay in [0, 1] (prep_slong!,
  src/storage_long.jl:77-146, commit_slong! :202-243);
* the level move (L = T) raises every instant and changes nothing else; its ΔC(t) is exactly
  zero while every instant stays in [0, 1] (mv_eval! :148-188) and the exact global change there
  is zero up to round-off;
* ΔC(t) equals the change of the global energy after the commit, also with states of charge
  outside [0, 1] and plateaus that wrap around the year (test/test_local.jl:120); the caches
  equal a rebuild (:122-125);
* 40000 draws against the quadrature CDF of exp(−β ΔC(t)), KS < 0.015 (:126-171);
* a whole long-range sweep (storage_long_sweep!, :245-298) makes one level move and `nmoves`
  plateaus per unit with at least `hmin` hours, and runs with its probability.
"""
import numpy as np
import pytest

import oracle
from esmsampling import synth

H2 = 1


def _state(seed=3, N=8, L=12, T=40, positive=False):
    s = synth.make_system(N=N, L=L, T=T, dt=1.0, seed=seed)
    R = 2
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    v[0] = rng.normal(0, 400.0, (s.N, s.T, R))
    v[1] = rng.uniform(0.2, 0.8, v[1].shape)
    v[2] = rng.uniform(0.2, 0.8, v[2].shape)
    v[3] = rng.uniform(0.2, 0.8, v[3].shape)
    if positive:                 # little renewable output, gas at 20 %, no flows, flat storage:
        x[:] = 0.02              # every node needs supply (a positive requirement)
        v[1] = 0.2
        v[0] = 0.0
        v[2:] = 0.5
    return s, v, s.k0.copy(), x


@pytest.mark.parametrize("a,L,n", [(3, 5, 2), (30, 15, 5), (38, 1, 0), (0, 39, 7)])
def test_plateau_directions(a, L, n):
    s, v, k, x = _state(positive=True)
    o = oracle.Oracle(s)
    _, _, r, F = o.energy(v, k, x)
    assert np.all(r[:, :, 1] > 0)
    t = 0.05
    v2, r2, F2 = v.copy(), r.copy(), F.copy()
    o.slong_apply(v2, k, r2, F2, 1, H2, n, a, L, t)
    dv = v2 - v
    T = s.T
    raised = [(a + q) % T for q in range(1, L + 1)]
    exp_e = np.zeros_like(dv[2 + H2])
    exp_e[n, raised, 1] = t
    np.testing.assert_allclose(dv[2 + H2], exp_e, atol=1e-15)
    assert np.all(dv[2] == 0) and np.all(dv[0] == 0) and np.array_equal(F2, F)
    ta, tb = a % T, (a + L) % T
    moved = {tuple(z) for z in np.argwhere(dv[1] != 0)}
    assert moved <= {(n, ta, 1), (n, tb, 1)}
    kst = np.clip(k[4 + H2, n], s.klo[4 + H2, n], s.khi[4 + H2, n])
    ks = kst * s.stor_hours[H2, n] / s.dt
    kg = np.clip(k[3, n], s.klo[3, n], s.khi[3, n])
    assert dv[1, n, ta, 1] * kg == pytest.approx(ks * t, rel=1e-9)
    assert dv[1, n, tb, 1] * kg == pytest.approx(-ks * t, rel=1e-9)
    interior = [(a + q) % T for q in range(1, L)]
    dr = r2 - r
    if interior:
        assert np.all(dr[n, interior, 1] == 0)            # exactly: inside the box nothing changes
    others = np.ones(s.N, bool)
    others[n] = False
    assert np.all(dr[others] == 0) and np.all(dr[:, :, 0] == 0)


def test_level_move_inside_and_outside_the_box():
    s, v, k, x = _state(seed=5)
    o = oracle.Oracle(s)
    parts, _, r, F = o.energy(v, k, x)
    n, j = 4, 0
    y = v[2 + H2, n, :, j]
    lo, hi = -y.min(), 1.0 - y.max()                     # every instant stays in [0, 1]
    ts = np.linspace(lo, hi, 9)[1:-1]
    dc = o.slong_cost(v, k, r, F, j, H2, n, 0, s.T, ts)
    assert np.all(dc == 0.0)
    for t in ts[[0, 3, -1]]:
        v2, r2, F2 = v.copy(), r.copy(), F.copy()
        o.slong_apply(v2, k, r2, F2, j, H2, n, 0, s.T, t)
        np.testing.assert_allclose(v2[2 + H2, n, :, j] - y, t, atol=1e-15)
        d = v2 - v
        d[2 + H2, n, :, j] = 0
        assert np.all(d == 0) and np.array_equal(r2, r)
        p2, _, _, _ = o.energy(v2, k, x)
        assert abs(p2[j, 1] - parts[j, 1]) <= 1e-12 * abs(parts[j, 1])
    # outside: the instants that leave [0, 1] clamp, the injections of their steps change
    for t in (hi + 0.03, lo - 0.07):
        dc = o.slong_cost(v, k, r, F, j, H2, n, 0, s.T, [t])[0]
        v2, r2, F2 = v.copy(), r.copy(), F.copy()
        o.slong_apply(v2, k, r2, F2, j, H2, n, 0, s.T, t)
        p2, _, r_, _ = o.energy(v2, k, x)
        assert dc != 0
        assert abs((p2[j, 1] - parts[j, 1]) - dc) <= 1e-12 * abs(parts[j, 1]) + 1e-9 * abs(dc)
        assert np.max(np.abs(r2 - r_)) < 1e-6 * np.abs(r_).max()


@pytest.mark.parametrize("seed", [4, 6])
def test_incremental_cost_equals_global(seed):
    s, v, k, x = _state(seed=seed)
    v[3] = np.random.default_rng(seed).uniform(-0.05, 1.05, v[3].shape)   # also states outside the box
    o = oracle.Oracle(s)
    parts, _, r, F = o.energy(v, k, x)
    rng = np.random.default_rng(seed + 10)
    for trial in range(16):
        L = int(rng.integers(1, s.T)) if trial % 5 else s.T
        a = int(rng.integers(0, s.T))
        n, j = int(rng.integers(s.N)), int(rng.integers(2))
        t = float(rng.normal(0, 0.3))
        dc = o.slong_cost(v, k, r, F, j, H2, n, a, L, [t])[0]
        v2, r2, F2 = v.copy(), r.copy(), F.copy()
        o.slong_apply(v2, k, r2, F2, j, H2, n, a, L, t)
        p2, _, r_, F_ = o.energy(v2, k, x)
        op0, op1 = parts[j, 1], p2[j, 1]
        assert abs((op1 - op0) - dc) <= 1e-12 * abs(op0) + 1e-9 * abs(dc), (trial, a, L, t)
        assert np.array_equal(p2[1 - j], parts[1 - j])
        assert np.max(np.abs(r2 - r_)) < 1e-6 * np.abs(r_).max()
        assert np.max(np.abs(F2 - F_)) < 1e-6 * max(np.abs(F_).max(), 1.0)


@pytest.mark.parametrize("a,L", [(5, 9), (33, 20), (0, 40)])
def test_line_draws_against_quadrature(a, L):
    s, v, k, x = _state(seed=7)
    o = oracle.Oracle(s)
    _, _, r, F = o.energy(v, k, x)
    n = 3
    rng = np.random.default_rng(2)
    U = rng.random((40000, 2))
    ts, st = o.slong_draw_many(v, k, r, F, 0, H2, n, a, L, U[:, 0], U[:, 1])
    assert np.all(st == 0)
    lo, hi = np.quantile(ts, 0.0005), np.quantile(ts, 0.9995)
    grid = np.linspace(lo - (hi - lo), hi + (hi - lo), 20001)
    lp = -s.beta_v * o.slong_cost(v, k, r, F, 0, H2, n, a, L, grid)
    w = np.exp(lp - lp.max())
    cdf = np.cumsum(w) / w.sum()
    tss = np.sort(ts)
    qs = np.quantile(ts, np.arange(1, 100) / 100)
    femp = np.searchsorted(tss, qs, side="right") / len(ts)
    fq = cdf[np.clip(np.searchsorted(grid, qs, side="right") - 1, 0, len(grid) - 1)]
    assert np.max(np.abs(femp - fq)) < 0.015


def test_long_sweep_moves_rate_and_caches():
    s, v, k, x = _state(seed=9)
    o = oracle.Oracle(s)
    R = v.shape[-1]
    kw = dict(x=x, threads=2, extra=0.0, atoms=False, p_storage_ms=0.0)
    v2, r2, F2, st = o.sweeps(v, k, 1, key=(3, 3), counter=0, p_storage_long=1.0, long_moves=6, **kw)
    assert np.all(st[:11, 0] == 0)
    # the H2 units (168 h) qualify, the batteries (6 h) do not: 6 plateaus and one level move each
    assert st[11, 0] == 6 * s.N * R and st[12, 0] == s.N * R
    assert st[11, 1] > 0.5 * st[11, 0] and st[11, 2] == 0 and st[11, 3] == 0
    _, _, r_, F_ = o.energy(v2, k, x)
    assert np.max(np.abs(r2 - r_)) < 1e-6 * np.abs(s.load).max()
    assert np.max(np.abs(v2[2] - v[2])) == 0 and np.max(np.abs(v2[0] - v[0])) == 0
    _, _, _, st = o.sweeps(v, k, 1, key=(3, 3), counter=0, p_storage_long=1.0, long_moves=2, long_hours=0.0, **kw)
    assert st[11, 0] == 2 * 2 * s.N * R and st[12, 0] == 2 * s.N * R
    runs = 0
    for c in range(40):
        _, _, _, st = o.sweeps(v, k, 1, key=(3, 3), counter=c, p_storage_long=0.5, long_moves=1, **kw)
        runs += st[12, 0] > 0
    assert 8 <= runs <= 32
