"""Pins of the oracle's dispatch following (the invariants of test/test_replicas.jl:205-247):
* the gas unit's output c(k)·ω is unchanged by its own capacity change (rtol 1e-9);
* renewable and storage units keep their utilizations (x and the SoC fractions untouched) and
  their capacities land at k0·e^u with u = log(k'/k0) (rtol 1e-12);
* where the gas unit can absorb the injection change (inside its box before and after), the
  node residual is unchanged (< 1e-6 MW): the compensation is exact;
* the same inputs give bitwise the same capacities (they are shared by every replica).
"""
import numpy as np

import oracle
from esmsampling import synth


def _setup(seed=4, R=3):
    s = synth.make_system(N=8, L=12, T=36, dt=1.0, seed=seed)
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    v[1] = rng.uniform(0.1, 0.9, v[1].shape)
    v[2] = rng.uniform(0.1, 0.9, v[2].shape)
    v[3] = rng.uniform(0.1, 0.9, v[3].shape)
    k = s.k0.copy()
    knew = k * np.exp(rng.normal(0, 0.02, k.shape))
    return s, v, k, x, knew


def test_gas_output_fixed():
    s, v, k, x, knew = _setup()
    kn = k.copy()
    kn[3] = knew[3]                       # only the gas capacities move
    o = oracle.Oracle(s)
    v2, k2 = o.follow(v, k, x, kn)
    out0 = k[3][:, None, None] * np.clip(v[1], 0, 1)
    out1 = k2[3][:, None, None] * np.clip(v2[1], 0, 1)
    inside = (v2[1] > 0) & (v2[1] < 1)
    np.testing.assert_allclose(out1[inside], out0[inside], rtol=1e-9)
    np.testing.assert_allclose(k2[3], kn[3], rtol=1e-15)
    r0 = o.energy(v, k, x)[2]
    r1 = o.energy(v2, k2, x)[2]
    assert np.max(np.abs(r1 - r0)[np.broadcast_to(inside, r0.shape)]) < 1e-6


def test_kop_curve_keeps_utilizations_and_balances():
    s, v, k, x, knew = _setup(seed=6)
    kn = knew.copy()
    kn[3] = k[3]                          # gas fixed: only the kop curves move
    o = oracle.Oracle(s)
    v2, k2 = o.follow(v, k, x, kn)
    np.testing.assert_allclose(k2, kn, rtol=1e-12)
    assert np.array_equal(v2[[0, 2, 3]], v[[0, 2, 3]])   # angles and SoC fractions fixed
    r0 = o.energy(v, k, x)[2]
    r1 = o.energy(v2, k2, x)[2]
    # P1 = (Q − Δinj)₊ + e: where the node is not dumping (r ≥ 0, so Q = r + κω ≥ 0 all along)
    # and the gas unit stays inside its box, the node residual is unchanged
    ok = (v2[1] > 0) & (v2[1] < 1) & (v[1] > 0) & (v[1] < 1) & (r0 >= 0)
    assert ok.sum() > 50
    assert np.max(np.abs(r1 - r0)[ok]) < 1e-6
    # where it is dumping by more than every injection change, the gas unit does not move
    dump = (r0 + k[3][:, None, None] * np.clip(v[1], 0, 1)) < -5e4
    assert np.array_equal(v2[1][dump], v[1][dump])


def test_bitwise_repeatable():
    s, v, k, x, knew = _setup(seed=7)
    o = oracle.Oracle(s)
    a = o.follow(v, k, x, knew)
    b = o.follow(v[..., :1], k, x[..., :1], knew)
    assert np.array_equal(a[1], b[1])           # the capacities do not depend on the replicas
    assert np.array_equal(a[0][..., :1], b[0])
