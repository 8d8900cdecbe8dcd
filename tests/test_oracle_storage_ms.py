"""Pins of the oracle's storage multiscale sweep (CPU only).

* a plateau moves the state of charge of its units by exactly t at the instants b0+1 … b0+L and
  nothing else of v except the gas at the block ends; while the states of charge stay inside
  [0, 1] the residuals of the inner steps do not change, and the gas units at the block ends
  follow the nominal requirement change ±κs t (κs = k h/Δt) while they are inside their boxes
  and the requirement stays positive (prep_storage_ms!, src/flowmoves.jl:187-244);
* ΔC(t) equals the change of the global energy after the commit (test/test_local.jl:120), the
  caches equal a rebuild (:122-125);
* 40000 draws against the quadrature CDF of exp(−β ΔC(t)), KS < 0.015 (:126-171);
* a whole storage sweep runs 8 moves per block of every length L < T/2 below the offset limits.
"""
import numpy as np
import pytest

import oracle
from esmsampling import synth


def _state(seed=3, N=10, L=16, T=40, C=None):
    s = synth.make_system(N=N, L=L, T=T, dt=1.0, seed=seed, ncountry=C)
    R = 2
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    v[0] = rng.normal(0, 400.0, (s.N, s.T, R))
    v[1] = rng.uniform(0.2, 0.8, v[1].shape)
    v[2] = rng.uniform(0.2, 0.8, v[2].shape)
    v[3] = rng.uniform(0.2, 0.8, v[3].shape)
    return s, v, s.k0.copy(), x


def _group(s, c, si):
    return si, np.flatnonzero(s.country == c).astype(np.int32)


def test_plateau_directions():
    s, v, k, x = _state()
    x[:] = 0.02                                   # little renewable output, gas at 20 %, no flows and
    v[1] = 0.2                                    # flat storage: every node needs supply (a
    v[0] = 0.0                                    # positive requirement)
    v[2:] = 0.5
    o = oracle.Oracle(s)
    _, _, r, F = o.energy(v, k, x)
    assert np.all(r[:, :, 1] > 0)
    for (L, b0, si, nodes) in [(4, 3, 0, [2]), (8, 10, 1, [5]), (2, 30, 0, list(_group(s, 1, 0)[1])),
                               (8, 5, 1, list(_group(s, 0, 1)[1]))]:
        nodes = np.array(nodes, dtype=np.int32)
        t = 0.05
        v2, r2, F2 = v.copy(), r.copy(), F.copy()
        o.sms_apply(v2, k, r2, F2, 1, L, b0, si, nodes, t)
        dv = v2 - v
        # the state of charge: +t on b0+1 … b0+L of the units, nowhere else
        exp_e = np.zeros_like(dv[2 + si])
        exp_e[nodes[:, None], np.arange(b0 + 1, b0 + L + 1)[None, :], 1] = t
        np.testing.assert_allclose(dv[2 + si], exp_e, atol=1e-15)
        assert np.all(dv[2 + (1 - si)] == 0) and np.all(dv[0] == 0) and np.array_equal(F2, F)
        # gas: only at the block ends, following ±κs t
        moved = np.argwhere(dv[1] != 0)
        assert set(map(tuple, moved[:, 1:])) <= {(b0, 1), (b0 + L, 1)}
        for n in nodes:
            kst = np.clip(k[4 + si, n], s.klo[4 + si, n], s.khi[4 + si, n])
            ks = kst * s.stor_hours[si, n] / s.dt
            kg = np.clip(k[3, n], s.klo[3, n], s.khi[3, n])
            assert dv[1, n, b0, 1] * kg == pytest.approx(ks * t, rel=1e-9)
            assert dv[1, n, b0 + L, 1] * kg == pytest.approx(-ks * t, rel=1e-9)
        # inner steps: the residuals do not move
        dr = r2 - r
        assert np.max(np.abs(dr[:, b0 + 1:b0 + L, 1])) < 1e-9 * np.abs(r).max()
        assert np.all(dr[:, :, 0] == 0)


@pytest.mark.parametrize("seed,N,L,C", [(4, 10, 16, None), (5, 10, 16, None), (6, 40, 70, 2)])
def test_incremental_cost_equals_global(seed, N, L, C):
    """ΔC(t) of a move equals the change of the global energy; the last case moves groups of 20
    units (two countries of 20 nodes)."""
    s, v, k, x = _state(seed=seed, N=N, L=L, C=C)
    v[2] = np.random.default_rng(seed).uniform(-0.05, 1.05, v[2].shape)   # also states outside the box
    o = oracle.Oracle(s)
    parts, _, r, F = o.energy(v, k, x)
    rng = np.random.default_rng(seed + 10)
    for trial in range(14):
        L = int(rng.choice([2, 4, 8, 16]))
        b0 = int(rng.integers(0, s.T - L - 1))
        j = int(rng.integers(2))
        if rng.random() < 0.5:
            si, nodes = _group(s, int(rng.integers(s.ncountry)), int(rng.integers(2)))
        else:
            si, nodes = int(rng.integers(2)), np.array([int(rng.integers(s.N))], dtype=np.int32)
        t = float(rng.normal(0, 0.4))
        dc = o.sms_cost(v, k, r, F, j, L, b0, si, nodes, [t])[0]
        v2, r2, F2 = v.copy(), r.copy(), F.copy()
        o.sms_apply(v2, k, r2, F2, j, L, b0, si, nodes, t)
        p2, _, r_, F_ = o.energy(v2, k, x)
        op0, op1 = parts[j, 1], p2[j, 1]
        assert abs((op1 - op0) - dc) <= 1e-12 * abs(op0) + 1e-9 * abs(dc), (trial, L, b0, t)
        assert np.array_equal(p2[1 - j], parts[1 - j])
        assert np.max(np.abs(r2 - r_)) < 1e-6 * np.abs(r_).max()


@pytest.mark.parametrize("L,b0,grp", [(4, 6, False), (8, 12, True), (2, 0, False)])
def test_line_draws_against_quadrature(L, b0, grp):
    s, v, k, x = _state(seed=7)
    o = oracle.Oracle(s)
    _, _, r, F = o.energy(v, k, x)
    si, nodes = _group(s, 1, 1) if grp else (0, np.array([3], dtype=np.int32))
    rng = np.random.default_rng(2)
    U = rng.random((40000, 2))
    ts, st = o.sms_draw_many(v, k, r, F, 0, L, b0, si, nodes, U[:, 0], U[:, 1])
    assert np.all(st == 0)
    lo, hi = np.quantile(ts, 0.0005), np.quantile(ts, 0.9995)
    grid = np.linspace(lo - (hi - lo), hi + (hi - lo), 20001)
    lp = -s.beta_v * o.sms_cost(v, k, r, F, 0, L, b0, si, nodes, grid)
    w = np.exp(lp - lp.max())
    cdf = np.cumsum(w) / w.sum()
    tss = np.sort(ts)
    qs = np.quantile(ts, np.arange(1, 100) / 100)
    femp = np.searchsorted(tss, qs, side="right") / len(ts)
    fq = cdf[np.clip(np.searchsorted(grid, qs, side="right") - 1, 0, len(grid) - 1)]
    assert np.max(np.abs(femp - fq)) < 0.015


def test_storage_sweep_moves_and_caches():
    s, v, k, x = _state(seed=9, T=40)
    o = oracle.Oracle(s)
    R = v.shape[-1]
    v2, r2, F2, st = o.sweeps(v, k, 1, key=(3, 3), counter=0, x=x, threads=2, extra=0.0, atoms=False,
                              p_storage_ms=1.0)
    assert np.all(st[:10, 0] == 0)
    # blocks: L in {2, 4, 8, 16} (L < T/2 = 20), offsets o in [0, L), nb = (T − L − 1 − o)//L + 1
    nmax = sum(8 * ((s.T - L - 1) // L + 1) for L in (2, 4, 8, 16))
    nmin = sum(8 * ((s.T - L - 1 - (L - 1)) // L + 1) for L in (2, 4, 8, 16))
    assert nmin * R <= st[10, 0] <= nmax * R
    assert st[10, 1] > 0.5 * st[10, 0] and st[10, 2] == 0 and st[10, 3] == 0
    _, _, r_, F_ = o.energy(v2, k, x)
    assert np.max(np.abs(r2 - r_)) < 1e-6 * np.abs(s.load).max()
    # with probability 1/2 per sweep
    runs = 0
    for c in range(40):
        _, _, _, st = o.sweeps(v, k, 1, key=(3, 3), counter=c, x=x, threads=2, extra=0.0, atoms=False)
        runs += st[10, 0] > 0
    assert 8 <= runs <= 32


def _requirement(s, v, r, k, si, n, tau, j):
    """Q = r + κ clamp(y) + (nominal − actual injection of the storage unit) at (n, τ): the
    requirement that the pro-rata compensation tracks (src/local.jl:1817-1840)."""
    kg = np.clip(k[3, n], s.klo[3, n], s.khi[3, n])
    kst = np.clip(k[4 + si, n], s.klo[4 + si, n], s.khi[4 + si, n])
    h, ein, eout = s.stor_hours[si, n], s.stor_eta_in[si, n], s.stor_eta_out[si, n]
    e0, e1 = v[2 + si, n, tau, j], v[2 + si, n, (tau + 1) % s.T, j]
    d = h * (np.clip(e1, 0, 1) - np.clip(e0, 0, 1)) / s.dt
    actual = kst * (d / ein if d >= 0 else -min(-d * eout, 1.0))
    nominal = kst * h * (e1 - e0) / s.dt
    return r[n, tau, j] + kg * np.clip(v[1, n, tau, j], 0, 1) + (nominal - actual)


def test_compensation_switches_off_at_the_requirement():
    """Beyond t* = −Q/g the compensating gas unit at the block end stops following: its output
    changes by exactly −Q, with Q the requirement r + κ clamp(y) + (nominal − actual storage
    injection) of the state before the move (prorata! with soc_qadj, src/local.jl:1782-1840).
    The storage unit charges at that step, so the injection term is not zero."""
    s, v, k, x = _state(seed=11)
    x[:] = 0.02
    v[1] = 0.5
    v[0] = 0.0
    v[2:] = 0.5
    L, b0, si, n = 4, 6, 0, 3
    tau = b0 + L
    v[2 + si, n, tau + 1, 0] = 0.6                 # charging from τ to τ + 1
    kst = np.clip(k[4 + si, n], s.klo[4 + si, n], s.khi[4 + si, n])
    ks = kst * s.stor_hours[si, n] / s.dt
    t = 0.1
    Q0 = 0.5 * ks * t                              # the requirement reaches zero at t* = t/2
    o = oracle.Oracle(s)
    _, _, r, F = o.energy(v, k, x)
    s.load[n, tau] += Q0 - _requirement(s, v, r, k, si, n, tau, 0)
    o = oracle.Oracle(s)
    _, _, r, F = o.energy(v, k, x)
    assert _requirement(s, v, r, k, si, n, tau, 0) == pytest.approx(Q0, rel=1e-12)
    v2, r2, F2 = v.copy(), r.copy(), F.copy()
    o.sms_apply(v2, k, r2, F2, 0, L, b0, si, np.array([n], dtype=np.int32), t)
    kg = np.clip(k[3, n], s.klo[3, n], s.khi[3, n])
    assert 0 < v2[1, n, tau, 0] < 1
    assert (v2[1, n, tau, 0] - v[1, n, tau, 0]) * kg == pytest.approx(-Q0, rel=1e-9)
