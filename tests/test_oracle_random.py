"""Pins of the oracle's random families (CPU only).

* the templates: one xloc template per (node, neighbour), one xpair template per ordered pair,
  the cells of the three partition levels, level probabilities ∝ number of cells;
* the directions against what the network fixes: an xpair move keeps every nodal residual while
  both gas units stay inside their boxes (L Δθ = e_n − e_m exactly: the grounded re-routing); an
  xloc move keeps the residuals of its ball {n, m} and moves each leak node by g·t with
  g = −Σ b_e Δθ_i; xlocc compensates that leak at the leak nodes while their gas units are in
  their boxes and the requirement stays positive; a cell move changes each of its nodes'
  residuals by −κ_n t (test/test_local.jl:69-83 re-routing balance);
* ΔC(t) evaluated on the fly equals the change of the oracle's global energy after the commit
  (test/test_local.jl:120), the caches equal a rebuild (:122-125);
* 40000 draws of t against the quadrature CDF of exp(−β ΔC(t)), KS < 0.015 (:126-171);
* the family mix of whole sweeps against the weights of plain_families (src/local.jl:1912-1915).
"""
import math

import numpy as np
import pytest

import oracle
from esmsampling import synth

KINDS = {"cell": 0, "xloc": 1, "xlocc": 2, "xpair": 3}


def _state(seed=3, N=8, L=12, T=24):
    s = synth.make_system(N=N, L=L, T=T, dt=1.0, seed=seed)
    R = 2
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    v[0] = rng.normal(0, 400.0, (s.N, s.T, R))
    v[1] = np.clip(v[1] + rng.normal(0, 0.05, v[1].shape), -0.02, 1.02)
    v[2] = rng.uniform(0.0, 1.0, v[2].shape)
    v[3] = rng.uniform(0.0, 1.0, v[3].shape)
    return s, v, s.k0.copy(), x


def _nbrs(s):
    nb = [set() for _ in range(s.N)]
    for a, b in zip(s.line_bus0, s.line_bus1):
        nb[int(a)].add(int(b))
        nb[int(b)].add(int(a))
    return nb


def test_template_layout():
    s, v, k, x = _state()
    o = oracle.Oracle(s)
    info = o.tables_info()
    nb = _nbrs(s)
    xp = info["xp_off"]
    assert xp[0] == 0 and all(xp[n + 1] - xp[n] == len(nb[n]) for n in range(s.N))
    assert info["xp0"] == xp[-1] and info["cl0"] == info["xp0"] + s.N * s.N
    assert info["cells"] == [1, s.ncountry, s.N]
    assert info["ntpl"] == info["cl0"] + 1 + s.ncountry + s.N
    # a cell of the country level holds exactly that country's nodes
    for c in range(s.ncountry):
        nodes, kk, _ = o.template_nodes(info["cl0"] + 1 + c)
        assert sorted(nodes) == sorted(np.flatnonzero(s.country == c)) and len(kk) == 0


def _apply(o, v, k, r, F, j, kind, tid, tau, t):
    v2, r2, F2 = v.copy(), r.copy(), F.copy()
    o.sparse_apply(v2, k, r2, F2, j, kind, tid, tau, t)
    return v2, r2, F2


def _inside(v, nodes, tau, j, lo=0.05, hi=0.95):
    return all(lo < v[1, n, tau, j] < hi for n in nodes)


def test_xpair_keeps_every_residual():
    s, v, k, x = _state(seed=4)
    v[1] = np.clip(v[1], 0.2, 0.8)                 # gas units well inside their boxes
    o = oracle.Oracle(s)
    _, _, r, F = o.energy(v, k, x)
    info = o.tables_info()
    rng = np.random.default_rng(1)
    for _ in range(10):
        n, m = rng.choice(s.N, 2, replace=False)
        tid = info["xp0"] + n * s.N + m
        t = 0.05 * min(k[3, n], k[3, m])          # gas changes ±t/κ of at most 0.05
        v2, r2, F2 = _apply(o, v, k, r, F, 0, KINDS["xpair"], tid, 5, t)
        assert np.max(np.abs(r2[:, 5, 0] - r[:, 5, 0])) < 1e-9 * np.abs(r).max()
        assert v2[1, n, 5, 0] == pytest.approx(v[1, n, 5, 0] + t / k[3, n], rel=1e-12)
        assert v2[1, m, 5, 0] == pytest.approx(v[1, m, 5, 0] - t / k[3, m], rel=1e-12)
        # the flows moved and agree with the angles (F = b (θ_a − θ_b) changes consistently)
        dF = F2[:, 5, 0] - F[:, 5, 0]
        dth = v2[0, :, 5, 0] - v[0, :, 5, 0]
        np.testing.assert_allclose(dF, s.line_b * (dth[s.line_bus0] - dth[s.line_bus1]), atol=1e-9 * abs(t))
        assert np.abs(dF).max() > 0


@pytest.mark.parametrize("kind", ["xloc", "xlocc"])
def test_xloc_ball_and_leak(kind):
    s, v, k, x = _state(seed=6)
    v[1] = np.clip(v[1], 0.2, 0.8)
    o = oracle.Oracle(s)
    _, _, r, F = o.energy(v, k, x)
    r[:, 7, 1] = np.abs(r[:, 7, 1]) + 1.0         # positive requirement everywhere (see below)
    info = o.tables_info()
    nb = _nbrs(s)
    checked = 0
    for n in range(s.N):
        for p, m in enumerate(sorted(nb[n])):
            tid = int(info["xp_off"][n] + p)
            nodes, leak, g = o.template_nodes(tid)
            assert {n, m} <= set(nodes.tolist()) and not ({n, m} & set(leak.tolist()))
            # leak nodes are the first layer outside the ball
            assert set(leak.tolist()) <= (nb[n] | nb[m]) - {n, m}
            t = 0.02 * min(k[3, n], k[3, m])
            v2, r2, F2 = _apply(o, v, k, r, F, 1, KINDS[kind], tid, 7, t)
            dr = r2[:, 7, 1] - r[:, 7, 1]
            tol = 1e-9 * np.abs(r[:, 7, 1]).max()
            assert abs(dr[n]) < tol and abs(dr[m]) < tol
            for q, lk in enumerate(leak):
                if kind == "xloc":
                    assert dr[lk] == pytest.approx(g[q] * t, rel=1e-9, abs=tol)
                elif 0.05 < v2[1, lk, 7, 1] < 0.95:
                    assert abs(dr[lk]) < tol          # compensated inside the box while Q > 0
            others = set(range(s.N)) - set(nodes.tolist())
            assert all(dr[i] == 0.0 for i in others)
            checked += 1
    assert checked == sum(len(a) for a in nb)


def test_cell_moves_every_gas_unit_of_the_cell():
    s, v, k, x = _state(seed=7)
    v[1] = np.clip(v[1], 0.2, 0.8)
    o = oracle.Oracle(s)
    _, _, r, F = o.energy(v, k, x)
    info = o.tables_info()
    for tid in range(info["cl0"], info["ntpl"]):
        nodes, _, _ = o.template_nodes(tid)
        v2, r2, F2 = _apply(o, v, k, r, F, 0, KINDS["cell"], tid, 3, 0.1)
        for n in range(s.N):
            if n in nodes:
                assert v2[1, n, 3, 0] == pytest.approx(v[1, n, 3, 0] + 0.1, rel=1e-14)
                assert r2[n, 3, 0] - r[n, 3, 0] == pytest.approx(-0.1 * k[3, n], rel=1e-9)
            else:
                assert v2[1, n, 3, 0] == v[1, n, 3, 0] and r2[n, 3, 0] == r[n, 3, 0]
        assert np.array_equal(F2, F) and np.array_equal(v2[0], v[0])


@pytest.mark.parametrize("kind", list(KINDS))
def test_incremental_cost_equals_global(kind):
    s, v, k, x = _state(seed=8)
    o = oracle.Oracle(s)
    parts, _, r, F = o.energy(v, k, x)
    info = o.tables_info()
    rng = np.random.default_rng(11)
    for trial in range(16):
        j = int(rng.integers(2))
        tau = int(rng.integers(s.T))
        if kind == "cell":
            tid = int(rng.integers(info["cl0"], info["ntpl"]))
            t = float(rng.normal(0, 0.3))
        elif kind == "xpair":
            n, m = rng.choice(s.N, 2, replace=False)
            tid = info["xp0"] + int(n) * s.N + int(m)
            t = float(rng.normal(0, 0.5)) * k[3].mean()
        else:
            tid = int(rng.integers(0, info["xp0"]))
            t = float(rng.normal(0, 0.5)) * k[3].mean()
        dc = o.sparse_cost(v, k, r, F, j, KINDS[kind], tid, tau, [t])[0]
        v2, r2, F2 = _apply(o, v, k, r, F, j, KINDS[kind], tid, tau, t)
        p2, _, r_, F_ = o.energy(v2, k, x)
        op0, op1 = parts[j, 1], p2[j, 1]
        assert abs((op1 - op0) - dc) <= 1e-12 * abs(op0) + 1e-9 * abs(dc), (trial, tid, t)
        assert np.array_equal(p2[1 - j], parts[1 - j])
        assert np.max(np.abs(r2 - r_)) < 1e-6 * np.abs(r_).max() and np.max(np.abs(F2 - F_)) < 1e-6 * max(np.abs(F_).max(), 1)


@pytest.mark.parametrize("kind,pick", [("cell", 0), ("cell", 5), ("xloc", 2), ("xlocc", 3), ("xlocc", 9),
                                       ("xpair", 17)])
def test_line_draws_against_quadrature(kind, pick):
    s, v, k, x = _state(seed=5)
    o = oracle.Oracle(s)
    _, _, r, F = o.energy(v, k, x)
    info = o.tables_info()
    if kind == "cell":
        tid = info["cl0"] + pick
    elif kind == "xpair":
        tid = info["xp0"] + (pick // s.N) * s.N + (pick % s.N)
    else:
        tid = pick
    j, tau = 0, 7
    rng = np.random.default_rng(9)
    U = rng.random((40000, 2))
    ts, st = o.sparse_draw_many(v, k, r, F, j, KINDS[kind], tid, tau, U[:, 0], U[:, 1])
    assert np.all(st == 0)
    lo, hi = np.quantile(ts, 0.0005), np.quantile(ts, 0.9995)
    grid = np.linspace(lo - (hi - lo), hi + (hi - lo), 20001)
    lp = -s.beta_v * o.sparse_cost(v, k, r, F, j, KINDS[kind], tid, tau, grid)
    w = np.exp(lp - lp.max())
    cdf = np.cumsum(w) / w.sum()
    tss = np.sort(ts)
    qs = np.quantile(ts, np.arange(1, 100) / 100)
    femp = np.searchsorted(tss, qs, side="right") / len(ts)
    fq = cdf[np.clip(np.searchsorted(grid, qs, side="right") - 1, 0, len(grid) - 1)]
    assert np.max(np.abs(femp - fq)) < 0.015


def test_family_mix_of_sweeps():
    """Attempted moves per family against round(4N) draws per step over the plain families
    (pair and cellpair are void: one dispatchable unit per node); every drawn unit, potential,
    cell and xloc / xlocc / xpair move of this network is valid."""
    s = synth.make_system(N=10, L=16, T=48, dt=1.0, seed=8)
    R = 2
    rng = np.random.default_rng(3)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    o = oracle.Oracle(s)
    sweeps = 3
    _, _, _, st = o.sweeps(v, s.k0, sweeps, key=(7, 7), counter=0, x=x, threads=2, atoms=False)
    assert np.all(st[:3, 0] == 0)                  # no atoms
    ndraw = round(4 * s.N) * s.T * sweeps          # per replica
    w = np.array([0.3, 0.25, 0.03, 0.08, 0.04, 0.1, 0.1, 0.1, 0.3])
    p = w / w.sum()
    # family rows: spair 3, unit 4, potential 5, cell 6, xloc 7, xlocc 8, xpair 9
    for row, wi in ((4, 0), (5, 5), (6, 7), (7, 4), (8, 3)):
        mean = ndraw * p[wi]
        sd = math.sqrt(ndraw * p[wi] * (1 - p[wi]))
        assert abs(st[row, 0] / R - mean) < 5 * sd, (row, st[row, 0] / R, mean)
    # xpair: the m = n draws are void (probability 1/N)
    mean = ndraw * p[2] * (1 - 1 / s.N)
    assert abs(st[9, 0] / R - mean) < 5 * math.sqrt(mean)
    # spair: valid where κs ≥ 1 and κ ≥ 1 (all nodes here)
    assert abs(st[3, 0] / R - ndraw * p[8]) < 5 * math.sqrt(ndraw * p[8])
    assert np.all(st[3:13, 1] > 0.3 * st[3:13, 0])


def test_sweeps_with_random_families_keep_caches():
    s = synth.make_system(N=10, L=16, T=47, dt=1.0, seed=2)   # odd T: three parity classes
    R = 2
    rng = np.random.default_rng(4)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    o = oracle.Oracle(s)
    v2, r2, F2, st = o.sweeps(v, s.k0, 2, key=(1, 9), counter=3, x=x, threads=2, p_xpair=0.1, xlong_moves=8)
    _, _, r_, F_ = o.energy(v2, s.k0, x)
    assert np.max(np.abs(r2 - r_)) < 1e-6 * np.abs(s.load).max()
    assert np.max(np.abs(F2 - F_)) < 1e-6 * max(np.abs(F_).max(), 1)
    assert np.all(st[:, 0] > 0) and np.all(st[:, 3] == 0)
