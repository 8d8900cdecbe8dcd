"""Pins of the oracle's cross-node storage pairs (CPU only; src/storage_cross.jl).

* the templates (sx_templates :29-53 with xloc_directions src/local.jl:122-191 at partner hops 1
  and ball radius 1): the angles of slot (n, m) move a unit of flow out of n and into m, nothing
  out of the other nodes of the ball, and the requirement change of each leak node equals the
  flow the template sends into it; the angle nodes lie in the ball around the pair; the exact
  global re-routing equals the Laplacian pseudo-inverse (numpy) when the ball covers the
  component; the footprint holds the pair, the angle nodes and their lines' ends, and the leaks;
* the partner law (sx_partner :62-90): over a fine grid of uniforms the partners come with
  frequencies ∝ min(k_A, k_B) among the units of the neighbourhood with hours ≥ hmin and κ ≥ 1,
  enumerated here from the model;
* a pair move (sxpair: L = 1; sxlong: L on 1 … T−1, wrapping) moves A's instants a+1 … a+L by t
  and B's by −ρ t, ρ = E_A/E_B, the angles by ∓κ_A t times the template at the end steps, and
  nothing of v else but the gas of the leak nodes at the end steps; its ΔC(t) equals the change of
  the global energy (also with states of charge outside [0, 1]), the caches equal a rebuild
  (test/test_local.jl:120-125);
* 40000 draws against the quadrature CDF of exp(−β ΔC(t)), KS < 0.015;
* a sweep with the pairs on makes the drawn number of moves: round(4·xmul·N) draws per step
  with sxpair's weight 0.1 after the plain families (xmul = 1.4/1.3, local_sweep! :1934-1941),
  and `nmoves` interval pairs per unit of the pairs when the long-range sweep runs.
"""
import math

import numpy as np
import pytest

import oracle
from esmsampling import synth

H2 = 1


def _sys(N=12, L=18, T=40, seed=21):
    return synth.make_system(N=N, L=L, T=T, dt=1.0, seed=seed)


def _state(s, seed, R=2, outside=False):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    v[0] = rng.normal(0, 300.0, (s.N, s.T, R))
    v[1] = rng.uniform(0.1, 0.9, v[1].shape)
    v[2:] = rng.uniform(-0.05, 1.05, v[2:].shape) if outside else rng.uniform(0.1, 0.9, v[2:].shape)
    return v, s.k0.copy(), x


def _nbrs(s):
    nb = [set() for _ in range(s.N)]
    for a, b in zip(s.line_bus0, s.line_bus1):
        if a != b:
            nb[a].add(int(b))
            nb[b].add(int(a))
    return nb


def _slots(o, n):
    out, k = [], 0
    while (sl := o.sx_slot(n, k)) is not None:
        out.append(sl)
        k += 1
    return out


def _outflow(s, th):
    """Net flow out of every node for node angles th (F = b (θ_a − θ_b) on each line)."""
    out = np.zeros(s.N)
    for e in range(s.L):
        a, b = s.line_bus0[e], s.line_bus1[e]
        f = s.line_b[e] * (th[a] - th[b])
        out[a] += f
        out[b] -= f
    return out


def test_templates_route_a_unit_dipole():
    s = _sys()
    o = oracle.Oracle(s)
    nb = _nbrs(s)
    nleak = nexact = 0
    for n in range(s.N):
        slots = _slots(o, n)
        assert slots[0]["partner"] == n and len(slots[0]["th"][0]) == 0 and len(slots[0]["leak"][0]) == 0
        assert [sl["partner"] for sl in slots[1:]] == sorted(nb[n])
        for sl in slots[1:]:
            m = sl["partner"]
            th = np.zeros(s.N)
            th[sl["th"][0]] = sl["th"][1]
            ball = {n, m} | nb[n] | nb[m]
            assert set(sl["th"][0]) <= ball
            out = _outflow(s, th)
            lk = dict(zip(sl["leak"][0].tolist(), sl["leak"][1].tolist()))
            assert set(lk) <= set(range(s.N)) - ball
            for i in range(s.N):
                want = (i == n) - (i == m) if i in ball else lk.get(i, 0.0)
                assert out[i] == pytest.approx(want, abs=1e-12), (n, m, i)
            if lk:
                nleak += 1
            else:
                nexact += 1
            foot = {n, m} | set(lk)
            for i in sl["th"][0]:
                foot.add(int(i))
                for e in range(s.L):
                    if i in (s.line_bus0[e], s.line_bus1[e]):
                        foot |= {int(s.line_bus0[e]), int(s.line_bus1[e])}
            assert set(sl["foot"].tolist()) == foot
    assert nleak > 0


def test_template_of_a_covering_ball_is_the_global_rerouting():
    """A 4-node chain: every ball of radius one around a pair of neighbours covers it or not; the
    covering ones equal the grounded Laplacian solve (numpy pinv, shifted to the ground node)."""
    s = synth.make_system(N=5, L=4, T=8, dt=1.0, seed=2)
    o = oracle.Oracle(s)
    nb = _nbrs(s)
    Lap = np.zeros((s.N, s.N))
    for e in range(s.L):
        a, b, w = s.line_bus0[e], s.line_bus1[e], s.line_b[e]
        Lap[a, a] += w; Lap[b, b] += w; Lap[a, b] -= w; Lap[b, a] -= w
    checked = 0
    for n in range(s.N):
        for sl in _slots(o, n)[1:]:
            m = sl["partner"]
            if len({n, m} | nb[n] | nb[m]) < s.N:
                continue
            rhs = np.zeros(s.N)
            rhs[n], rhs[m] = 1.0, -1.0
            x = np.linalg.pinv(Lap) @ rhs
            x -= x[0]                                   # grounded at the component's first node
            th = np.zeros(s.N)
            th[sl["th"][0]] = sl["th"][1]
            np.testing.assert_allclose(th, np.where(np.abs(x) >= 1e-15, x, 0.0), atol=1e-12)
            assert len(sl["leak"][0]) == 0
            checked += 1
    assert checked > 0


@pytest.mark.parametrize("hmin", [24.0, 0.0])
def test_partner_law(hmin):
    s = _sys()
    o = oracle.Oracle(s)
    k = s.k0.copy()
    rng = np.random.default_rng(1)
    k[4:] *= rng.uniform(0.3, 3.0, k[4:].shape)       # distinct storage capacities
    kc = np.clip(k, s.klo, s.khi)
    nb = _nbrs(s)
    G = 4000
    for (sa, na) in [(H2, 0), (H2, 5), (0, 3)]:
        kA = kc[4 + sa, na]
        cands = []
        for mm in [na] + sorted(nb[na]):
            for si in range(2):
                if (si, mm) == (sa, na) or s.stor_hours[si, mm] < hmin:
                    continue
                kB = kc[4 + si, mm]
                if kB * s.stor_hours[si, mm] / s.dt >= 1.0:
                    cands.append(((si, mm), min(kA, kB)))
        got = {}
        for u in (np.arange(G) + 0.5) / G:
            p = o.sx_partner(k, sa, na, u, hmin=hmin)
            got[p] = got.get(p, 0) + 1
        if not cands:
            assert got == {None: G}
            continue
        W = sum(w for _, w in cands)
        assert set(got) == {c for c, _ in cands}
        for c, w in cands:
            assert abs(got[c] / G - w / W) <= 1.5 / G, (sa, na, c)


def _pick_pair(o, s, na, slot, sa=H2):
    sl = _slots(o, na)[slot]
    return (sa, na), (H2, sl["partner"]) if sl["partner"] != na else (1 - sa, na)


@pytest.mark.parametrize("seed", [3, 8])
def test_incremental_cost_equals_global(seed):
    s = _sys(seed=seed)
    o = oracle.Oracle(s)
    v, k, x = _state(s, seed, outside=True)
    parts, _, r, F = o.energy(v, k, x)
    rng = np.random.default_rng(seed + 1)
    for trial in range(20):
        na = int(rng.integers(s.N))
        slots = _slots(o, na)
        q = int(rng.integers(len(slots)))
        A, B = _pick_pair(o, s, na, q, sa=H2 if trial % 3 else 0)
        L = 1 if trial % 2 else int(rng.integers(1, s.T))
        a = int(rng.integers(0, s.T))
        W = 1 if L == 1 and trial % 4 == 1 else 64
        j = int(rng.integers(2))
        t = float(rng.normal(0, 0.2))
        dc = o.sx_cost(v, k, r, F, j, A, B, a, L, [t], W=W)[0]
        v2, r2, F2 = v.copy(), r.copy(), F.copy()
        o.sx_apply(v2, k, r2, F2, j, A, B, a, L, t, W=W)
        p2, _, r_, F_ = o.energy(v2, k, x)
        assert abs((p2[j, 1] - parts[j, 1]) - dc) <= 1e-12 * abs(parts[j, 1]) + 1e-9 * abs(dc), (trial, a, L, t)
        assert np.array_equal(p2[1 - j], parts[1 - j])
        assert np.max(np.abs(r2 - r_)) < 1e-6 * np.abs(r_).max()
        assert np.max(np.abs(F2 - F_)) < 1e-6 * max(np.abs(F_).max(), 1.0)


@pytest.mark.parametrize("a,L", [(6, 1), (30, 17), (35, 12)])
def test_move_directions(a, L):
    """What a pair move changes: A's and B's instants, the angles at the end steps, the gas of the
    leak nodes there, and nothing else of v."""
    s = _sys()
    o = oracle.Oracle(s)
    v, k, x = _state(s, 5)
    _, _, r, F = o.energy(v, k, x)
    na = 2
    sl = _slots(o, na)[1]
    A, B = (H2, na), (H2, sl["partner"])
    t = 0.01
    v2, r2, F2 = v.copy(), r.copy(), F.copy()
    o.sx_apply(v2, k, r2, F2, 1, A, B, a, L, t, W=64)
    d = v2 - v
    T = s.T
    kc = np.clip(k, s.klo, s.khi)
    EA, EB = kc[4 + H2, A[1]] * s.stor_hours[H2, A[1]], kc[4 + H2, B[1]] * s.stor_hours[H2, B[1]]
    raised = [(a + q) % T for q in range(1, L + 1)]
    np.testing.assert_allclose(d[2 + H2, A[1], raised, 1], t, rtol=1e-12)
    np.testing.assert_allclose(d[2 + H2, B[1], raised, 1], -EA / EB * t, rtol=1e-9)
    ta, tb = a % T, (a + L) % T
    kap = EA / s.dt
    th = np.zeros(s.N)
    th[sl["th"][0]] = sl["th"][1]
    for tau, c in ((ta, -kap), (tb, kap)):
        np.testing.assert_allclose(d[0, :, tau, 1], np.where(np.abs(c * th) >= 1e-14, c * th, 0.0) * t, rtol=1e-9,
                                   atol=1e-12)
    mask = np.ones(d.shape, bool)
    mask[2 + H2, A[1], raised, 1] = mask[2 + H2, B[1], raised, 1] = False
    mask[0, :, [ta, tb], 1] = False
    leaks = sl["leak"][0]
    for tau in (ta, tb):
        mask[1, leaks, tau, 1] = False
    assert np.all(d[mask] == 0)


@pytest.mark.parametrize("a,L,W", [(5, 1, 1), (20, 9, 64), (36, 30, 64)])
def test_line_draws_against_quadrature(a, L, W):
    s = _sys()
    o = oracle.Oracle(s)
    v, k, x = _state(s, 7)
    _, _, r, F = o.energy(v, k, x)
    A, B = _pick_pair(o, s, 4, 1)
    rng = np.random.default_rng(2)
    U = rng.random((40000, 2))
    ts, st = o.sx_draw_many(v, k, r, F, 0, A, B, a, L, U[:, 0], U[:, 1], W=W)
    assert np.all(st == 0)
    lo, hi = np.quantile(ts, 0.0005), np.quantile(ts, 0.9995)
    grid = np.linspace(lo - (hi - lo), hi + (hi - lo), 20001)
    lp = -s.beta_v * o.sx_cost(v, k, r, F, 0, A, B, a, L, grid, W=W)
    w = np.exp(lp - lp.max())
    cdf = np.cumsum(w) / w.sum()
    tss = np.sort(ts)
    qs = np.quantile(ts, np.arange(1, 100) / 100)
    femp = np.searchsorted(tss, qs, side="right") / len(ts)
    fq = cdf[np.clip(np.searchsorted(grid, qs, side="right") - 1, 0, len(grid) - 1)]
    assert np.max(np.abs(femp - fq)) < 0.015


def test_sweep_move_counts():
    s = _sys(T=24)
    o = oracle.Oracle(s)
    v, k, x = _state(s, 9)
    R = v.shape[-1]
    kw = dict(x=x, threads=2, atoms=False, p_storage_ms=0.0)
    # sxpair: round(4·(1.4/1.3)·N) draws per step, a share 0.1/1.4 of them storage pairs; every
    # drawn H2 unit has partners here
    _, _, _, st = o.sweeps(v, k, 2, key=(2, 2), counter=0, p_storage_long=0.0, p_xpair=0.1, xlong_moves=0, **kw)
    ndraw = round(4 * (1.4 / 1.3) * s.N) * s.T * 2
    p = 0.1 / 1.4
    assert abs(st[13, 0] / R - ndraw * p) < 5 * math.sqrt(ndraw * p * (1 - p))
    w = np.array([0.3, 0.25, 0.03, 0.08, 0.04, 0.1, 0.1, 0.1, 0.3])
    assert abs(st[4, 0] / R - ndraw * 0.3 / 1.4) < 5 * math.sqrt(ndraw * 0.3 / 1.4)   # unit keeps its rate
    # sxlong: 5 pairs per H2 unit whenever the long-range sweep runs (always, p = 1)
    _, _, _, st = o.sweeps(v, k, 1, key=(2, 2), counter=0, extra=0.0, p_storage_long=1.0, long_moves=0,
                           p_xpair=0.0, xlong_moves=5, **kw)
    assert st[14, 0] == 5 * s.N * R and np.all(st[:11, 0] == 0) and st[14, 2] == 0
    runs = 0
    for c in range(30):
        _, _, _, st = o.sweeps(v, k, 1, key=(2, 2), counter=c, extra=0.0, p_storage_long=0.5, long_moves=0,
                               xlong_moves=1, **kw)
        runs += st[14, 0] > 0
    assert 5 <= runs <= 25
