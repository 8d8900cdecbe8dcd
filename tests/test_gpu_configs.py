"""GPU checks at the shapes of the benchmark configurations that the unit tests do not reach.

* the 256-node / 450-line network (uneven batch, full size): FP64 full sweeps (atoms, every
  random family, both storage sweeps) against the oracle at R = 3 and R = 64; FP32 energy,
  caches and scores against the oracle;
* a batch of uneven replica sets (40 and 256 nodes, different horizons and replica counts) swept
  as one captured graph, every set against the oracle's sweeps;
* the 3-hour resolution of the CPU-runnable configuration (dt = 3, one replica per chain);
* the heuristic start state set on the device against the host formula;
* the full-size configuration (256 nodes × 192,720 hourly steps × 32 replicas, FP32) through
  properties that do not need the oracle: the per-step costs sum to the totals, the caches
  equal a rebuild, everything finite, no improper lines or guard rejections, the moves as
  scheduled.
FP64 tolerances as in test_gpu_sweep.py: identical move statistics, states to 1e-9 of their
scale after whole sweeps.
"""
import numpy as np
import pytest

import oracle
from conftest import random_state
from esmsampling import FP32, FP64, api, synth

pytestmark = pytest.mark.gpu


def _start(s, R, seed, theta=300.0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    v[0] = rng.normal(0, theta, (s.N, s.T, R))
    return v, x


def _check_sweep(s, st, st_, vd, v_, rd, r_, Fd, F_):
    assert np.array_equal(st[:, :4], st_[:, :4]), (st[:, :4], st_[:, :4])
    np.testing.assert_allclose(vd[1:], v_[1:], atol=1e-9)
    np.testing.assert_allclose(vd[0], v_[0], atol=1e-9 * np.abs(v_[0]).max())
    np.testing.assert_allclose(rd, r_, atol=1e-9 * np.abs(s.load).max())
    np.testing.assert_allclose(Fd, F_, atol=1e-9 * np.abs(F_).max())


@pytest.mark.parametrize("R,T", [(3, 36), (64, 8)])
def test_full_sweep_fp64_256_nodes(R, T):
    """Atoms, every random family at the reference's rates and both storage sweeps on the
    256-node / 450-line network (the xpair templates span 450 lines and 256 nodes)."""
    print(f"full sweep fp64 256 nodes R={R} T={T}", flush=True)
    s = synth.make_system(N=256, L=450, T=T, dt=1.0, seed=4)
    v, x = _start(s, R, 1)
    dev = api.Replicas(s, R, FP64)
    dev.upload(v=v, k=s.k0, x=x)
    api.sweep_options(dev, p_storage_ms=1.0)
    st = api.operation_sweep(dev, 1, key=(2, 3), counter=1)
    vd, _, _ = dev.download()
    rd, Fd = dev.caches()
    print("  device sweep done; oracle", flush=True)
    v_, r_, F_, st_ = oracle.Oracle(s).sweeps(v, s.k0, 1, key=(2, 3), counter=1, x=x, threads=16,
                                               p_storage_ms=1.0)
    assert np.all(st_[:10, 0] > 0) and st_[11, 0] > 0
    _check_sweep(s, st, st_, vd, v_, rd, r_, Fd, F_)
    dev.close()


def test_energy_and_scores_fp32_256_nodes():
    print("energy and scores fp32 256 nodes", flush=True)
    s = synth.make_system(N=256, L=450, T=48, dt=1.0, seed=4)
    R = 4
    v, k, x = random_state(s, R, seed=3)
    v, x = (a.astype(np.float32).astype(np.float64) for a in (v, x))
    dev = api.Replicas(s, R, FP32)
    dev.upload(v=v, k=k, x=x)
    parts, ph = api.energy_cost(dev, per_hour=True)
    r, F = dev.caches()
    o = oracle.Oracle(s)
    parts_, ph_, r_, F_ = o.energy(v, k, x)
    tol = 1e-5
    np.testing.assert_allclose(F, F_, rtol=tol, atol=tol * np.abs(F_).max())
    np.testing.assert_allclose(r, r_, rtol=tol, atol=tol * np.abs(s.load).max())
    for p in (0, 1, 2, 3, 6, 7):
        np.testing.assert_allclose(parts[:, p], parts_[:, p], rtol=tol, err_msg=api.PART_NAMES[p])
    np.testing.assert_allclose(ph, ph_, rtol=tol, atol=tol * np.abs(ph_).max())
    for measure in ("utilization", "physical"):
        api.score_measure(dev, measure)
        sk, fk = api.capacity_scores(dev)
        sk_, fk_, sa = o.scores(v, k, x, measure=measure)
        np.testing.assert_allclose(fk, fk_, rtol=1e-14)
        assert np.all(np.abs(sk - sk_) <= tol * sa + 1e-300), measure
    dev.close()


def test_uneven_batch_with_256_node_set_matches_oracle():
    """Replica sets of 40 and 256 nodes, different horizons and replica counts, swept as one
    captured graph twice (the second replay after eager work on the sets): every set's states,
    caches and statistics against the oracle's sweeps under the set's key and counters."""
    print("uneven batch 40/256 nodes vs oracle", flush=True)
    shapes = [(40, 70, 24, 3, 1), (256, 450, 12, 2, 2), (40, 70, 30, 64, 3)]
    sets, starts = [], []
    for N, L, T, R, seed in shapes:
        s = synth.make_system(N=N, L=L, T=T, dt=1.0, seed=seed)
        v, x = _start(s, R, seed)
        d = api.Replicas(s, R, FP64)
        d.upload(v=v, k=s.k0, x=x)
        api.sweep_options(d, p_storage_ms=1.0)
        sets.append(d)
        starts.append((s, v, x))
    keys = [(9, i) for i in range(len(sets))]
    b = api.Batch(sets)
    st1 = b.sweeps(1, keys=keys, counters=[5, 6, 7])
    for d in sets:                               # eager work between the replays
        api.energy_cost(d)
    st2 = b.sweeps(1, keys=keys, counters=[8, 9, 10])
    b.close()
    for i, (d, (s, v, x)) in enumerate(zip(sets, starts)):
        print(f"  set {i}: oracle", flush=True)
        o = oracle.Oracle(s)
        v1, r1, F1, sa = o.sweeps(v, s.k0, 1, key=keys[i], counter=5 + i, x=x, threads=16, p_storage_ms=1.0)
        v2, r2, F2, sb = o.sweeps(v1, s.k0, 1, key=keys[i], counter=8 + i, x=x, r=r1, F=F1, threads=16,
                                  p_storage_ms=1.0)
        assert np.array_equal(st1[i][:, :4], sa[:, :4])
        vd, _, _ = d.download()
        rd, Fd = d.caches()
        _check_sweep(s, st2[i], sb, vd, v2, rd, r2, Fd, F2)
        d.close()


@pytest.mark.parametrize("R", [1, 4])
def test_three_hour_resolution_fp64(R):
    """The 3-hour resolution (dt = 3, the CPU-runnable configuration; one replica per chain):
    a full sweep with both storage sweeps against the oracle, and the energy with its per-step
    costs."""
    print(f"dt=3 sweep fp64 R={R}", flush=True)
    s = synth.make_system(N=40, L=70, T=64, dt=3.0, seed=1)
    v, x = _start(s, R, 6)
    dev = api.Replicas(s, R, FP64)
    dev.upload(v=v, k=s.k0, x=x)
    api.sweep_options(dev, p_storage_ms=1.0)
    st = api.operation_sweep(dev, 2, key=(4, 1), counter=3)
    vd, _, _ = dev.download()
    rd, Fd = dev.caches()
    o = oracle.Oracle(s)
    v_, r_, F_, st_ = o.sweeps(v, s.k0, 2, key=(4, 1), counter=3, x=x, threads=8, p_storage_ms=1.0)
    assert st_[10, 0] > 0 and st_[11, 0] > 0
    _check_sweep(s, st, st_, vd, v_, rd, r_, Fd, F_)
    parts, ph = api.energy_cost(dev, per_hour=True)
    parts_, ph_, _, _ = o.energy(v_, s.k0, x)
    np.testing.assert_allclose(parts[:, :4], parts_[:, :4], rtol=1e-9, atol=1e-9 * np.abs(parts_[:, 1]).max())
    np.testing.assert_allclose(ph.sum(axis=0), parts[:, 1], rtol=1e-12)
    dev.close()


@pytest.mark.parametrize("prec", [FP64, FP32])
def test_heuristic_state_on_the_device(prec):
    print(f"heuristic state prec={prec}", flush=True)
    s = synth.make_system(N=12, L=18, T=50, dt=1.0, seed=3)
    R = 5
    dev = api.Replicas(s, R, prec)
    api.sample_weather(dev, key=(3, 3))
    k = s.k0 * np.random.default_rng(1).uniform(0.5, 1.5, s.k0.shape)
    dev.upload(k=k)
    dev.heuristic()
    vd, _, x = dev.download()
    v_ = synth.heuristic_state(s, x, k)
    tol = 1e-13 if prec == FP64 else 1e-6
    np.testing.assert_allclose(vd, v_, atol=tol)
    _, _, drift = api.LocalCost(dev)
    assert np.all(drift == 0.0)
    dev.close()


@pytest.mark.timeout(900)
def test_full_size_256_nodes_192720_hours_32_replicas():
    """The full-size configuration (256 nodes / 450 lines, 22 hourly weather years = 192,720
    steps, 32 replicas, FP32: 1.6·10⁹ elements per node plane, 2.8·10⁹ per line field) on the
    device alone: prior weather, heuristic start, energy, one operation sweep with both storage
    sweeps and the reference configuration's storage pairs (sxpair 0.1, sxlong 8: hub templates
    in the 64-lane form with two kink and three balance slots), energy again. Checks: per-step costs sum to the totals; the incremental caches equal
    a rebuild (FP32 drift); all parts finite (a non-finite value anywhere in the state would
    reach them); no improper lines and no guard rejections; as many atoms and random moves as
    the schedule makes, every family accepting moves, at least one ΔC evaluation per move."""
    N, L, T, R = 256, 450, 192720, 32
    print(f"full size {N} nodes x {T} h x {R} replicas: system", flush=True)
    s = synth.make_system(N=N, L=L, T=T, dt=1.0, seed=5)
    dev = api.Replicas(s, R, FP32)
    print("  weather and heuristic start", flush=True)
    dev.upload(k=s.k0)
    api.sample_weather(dev, key=(5, 1))
    dev.heuristic()
    parts0, ph0 = api.energy_cost(dev, per_hour=True)
    assert np.all(np.isfinite(parts0)) and np.all(np.isfinite(ph0))
    np.testing.assert_allclose(ph0.sum(axis=0), parts0[:, 1], rtol=1e-9)
    print("  one operation sweep", flush=True)
    api.sweep_options(dev, p_storage_ms=1.0)
    api.storage_cross_options(dev, p_xpair=0.1, nxlong=8)
    st = api.operation_sweep(dev, 1, key=(5, 2), counter=0)
    print(f"  sweep statistics: {st[:, 0].tolist()}", flush=True)
    drift = api.rebuild_drift(dev)
    parts, ph = api.energy_cost(dev, per_hour=True)
    assert np.all(np.isfinite(parts)) and np.all(np.isfinite(ph))
    np.testing.assert_allclose(ph.sum(axis=0), parts[:, 1], rtol=1e-9)
    assert np.all(drift < 1e-2 * np.abs(s.load).max()), drift.max()
    # atoms: gas and angles at every (node, step), SoC for both carriers at every (node, step)
    assert st[0, 0] == N * T * R and st[1, 0] == N * T * R and st[2, 0] == 2 * N * T * R
    # random moves: round(4N) draws per step over the plain families (pair, cellpair void) and sxpair
    assert 0 < st[3:10, 0].sum() + st[13, 0] <= 4 * N * T * R * 1.2
    assert st[10, 0] > 0 and st[11, 0] > 0 and st[12, 0] > 0
    assert np.all(st[:, 2] == 0) and np.all(st[:, 3] == 0), st[:, 2:4]
    # the cross-node storage pairs (the last two families) made and accepted moves
    assert st[13, 0] > 0 and st[14, 0] > 0
    assert np.all(st[:, 1] > 0) and np.all(st[:, 4] >= st[:, 0])
    dev.close()
