"""Batched sweeps (one HIP graph over replica sets of uneven network, horizon, replica count and
precision) against the same sweeps run eagerly on twin sets, replay after replay; and the launch
autotuner: it leaves the chain where it was, and its choice gives the states of that launch form."""
import json

import numpy as np
import pytest

from esmsampling import FP32, FP64, api, synth

pytestmark = pytest.mark.gpu

SHAPES = [  # (N, L, T, R, precision)
    (6, 8, 10, 3, FP64),
    (5, 6, 40, 2, FP64),     # block lengths 2, 4, 8, 16: a new order of kernels every sweep
    (8, 12, 8, 64, FP32),
]


def _make(N, L, T, R, prec, seed):
    s = synth.make_system(N=N, L=L, T=T, dt=1.0, seed=seed)
    rng = np.random.default_rng(seed + 1)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    v[0] = rng.normal(0, 200.0, (s.N, s.T, R))
    pair = []
    for _ in range(2):
        d = api.Replicas(s, R, prec)
        d.upload(v=v, k=s.k0, x=x)
        pair.append(d)
    return pair


def _same(a, b):
    va, _, _ = a.download()
    vb, _, _ = b.download()
    ra, Fa = a.caches()
    rb, Fb = b.caches()
    return np.array_equal(va, vb) and np.array_equal(ra, rb) and np.array_equal(Fa, Fb)


def test_batch_graph_matches_eager_sweeps():
    print("batch graph vs eager, 3 uneven sets", flush=True)
    sets = [_make(*sh, seed=11 + i) for i, sh in enumerate(SHAPES)]
    batched = [p[0] for p in sets]
    eager = [p[1] for p in sets]
    b = api.Batch(batched)
    keys = [(3, 4), (5, 6), (7, 8)]
    # replays 0-3: every storage sweep on, whose block lengths run in a new order each sweep (the
    # lengths and their order drawn on the device: one instantiation, then updates in place);
    # 4-6: the multiscale sweep off, another kernel sequence (instantiated once more, then updated)
    plan = [(1.0, c) for c in range(4)] + [(0.0, c) for c in range(4, 7)]
    for rep, (pms, ctr) in enumerate(plan):
        for d in batched + eager:
            api.sweep_options(d, p_storage_ms=pms)
        counters = [10 * ctr + i for i in range(3)]
        st = b.sweeps(1, keys=keys, counters=counters)
        for i, d in enumerate(eager):
            st_e = api.operation_sweep(d, 1, key=keys[i], counter=counters[i])
            assert np.array_equal(st[i], st_e), f"replay {rep}, set {i}: statistics"
            assert _same(batched[i], d), f"replay {rep}, set {i}: states or caches"
        # eager work of the program between the replays (on both twins: the energy rebuilds the
        # caches of set 1 from the state)
        api.energy_cost(batched[1])
        api.energy_cost(eager[1])
        batched[2].download()
    launches, inst, upd = b.counts()
    assert launches == len(plan) and inst + upd == launches
    assert (inst, upd) == (2, 5), (inst, upd)
    b.close()
    for d in batched + eager:
        d.close()


def test_batch_graph_with_storage_sweeps_is_instantiated_once():
    """Ten replays of two sweeps each with both storage sweeps on at probability 1/2 (whether each
    runs, its block lengths, their order, offsets and classes are new draws every sweep) and the
    storage pairs on (the sxlong list and its rounds are drawn on the device): one instantiation,
    nine in-place updates, and every replay bitwise the eager sweeps' states, caches and
    statistics."""
    print("batch graph, storage sweeps at p = 1/2, 10 replays", flush=True)
    sets = [_make(*sh, seed=41 + i) for i, sh in enumerate(SHAPES)]
    batched = [p[0] for p in sets]
    eager = [p[1] for p in sets]
    for d in batched + eager:
        api.sweep_options(d, p_storage_ms=0.5)
        api.storage_long_options(d, p=0.5, nmoves=8, hmin=1.0)
        api.storage_cross_options(d, p_xpair=0.1, nxlong=4)
    b = api.Batch(batched)
    keys = [(13, 4), (15, 6), (17, 8)]
    ran = np.zeros(4, dtype=np.int64)
    for rep in range(10):
        counters = [100 * rep + 7 * i for i in range(3)]
        st = b.sweeps(2, keys=keys, counters=counters)
        for i, d in enumerate(eager):
            st_e = api.operation_sweep(d, 2, key=keys[i], counter=counters[i])
            assert np.array_equal(st[i], st_e), f"replay {rep}, set {i}: statistics"
            assert _same(batched[i], d), f"replay {rep}, set {i}: states or caches"
        ran += [st[1][api.FAMILY_NAMES.index(f), 0] > 0 for f in ("storage_ms", "slong", "sxpair", "sxlong")]
    launches, inst, upd = b.counts()
    assert (launches, inst, upd) == (10, 1, 9), (launches, inst, upd)
    # both storage sweeps and both pair families ran in some replays of set 1
    assert ran.min() > 0, ran
    b.close()
    for d in batched + eager:
        d.close()


def test_batch_rejects_bad_use():
    print("batch argument checks", flush=True)
    (a, a2), = [_make(5, 6, 8, 2, FP64, seed=3)]
    with pytest.raises(Exception, match="twice"):
        api.Batch([a, a])
    b = api.Batch([a, a2])
    with pytest.raises(ValueError):
        b.sweeps(1, keys=[(1, 1)], counters=[0, 0])
    a.lib.esm_profile(a.handle, 1)
    with pytest.raises(Exception, match="profiling"):
        b.sweeps(1)
    a.lib.esm_profile(a.handle, 0)
    b.sweeps(1)
    b.close()
    a.close()
    a2.close()


@pytest.mark.parametrize("R,prec", [(64, FP32), (3, FP64)])
def test_autotune_keeps_the_chain_and_its_choice_reproduces(R, prec, tmp_path):
    print(f"autotune R={R} prec={prec}", flush=True)
    d, twin = _make(8, 12, 8, R, prec, seed=21)
    api.sweep_options(d, p_storage_ms=1.0)
    api.sweep_options(twin, p_storage_ms=1.0)
    v0, _, _ = d.download()
    r0, F0 = d.caches()
    cache = str(tmp_path / "tune.json")
    res = api.autotune(d, nsweeps=1, key=(9, 9), counter=5, cache=cache)
    assert list(res["ms"]) == api.autotune_candidates(R) and all(t > 0 for t in res["ms"].values())
    assert (res["form"], res["block"], res["pad"]) == min(res["ms"], key=res["ms"].get)
    v1, _, _ = d.download()
    r1, F1 = d.caches()
    assert np.array_equal(v0, v1) and np.array_equal(r0, r1) and np.array_equal(F0, F1)
    # the twin runs the chosen form with the default block and the other padding: neither changes
    # results
    api.sweep_launch(twin, res["form"])
    api.sweep_pad(twin, not res["pad"])
    st = api.operation_sweep(d, 1, key=(1, 2), counter=0)
    st_t = api.operation_sweep(twin, 1, key=(1, 2), counter=0)
    assert np.array_equal(st, st_t) and _same(d, twin)
    # the cache gives the same choice without timing
    again = api.autotune(twin, cache=cache)
    assert again["cached"] and (again["form"], again["block"], again["pad"]) == (res["form"], res["block"], res["pad"])
    assert list(json.load(open(cache))) == [api._shape_key(d)]
    d.close()
    twin.close()


@pytest.mark.parametrize("R,prec", [(3, FP64), (16, FP32), (70, FP64)])
def test_padded_steps_change_nothing(R, prec):
    """Whole waves per step (R rounded up to a multiple of 64) against R threads per step, merged
    and per family: bitwise the same states, caches and statistics."""
    print(f"padded random launches R={R} prec={prec}", flush=True)
    a, b = _make(7, 10, 12, R, prec, seed=31)
    for merged in (True, False):
        for d in (a, b):
            api.sweep_options(d, p_storage_ms=1.0)
            api.sweep_launch(d, merged)
        api.sweep_pad(a, False)
        api.sweep_pad(b, True)
        api.sweep_block(b, 64)
        for c in range(2):
            sa = api.operation_sweep(a, 1, key=(4, 5), counter=c + (10 if merged else 0))
            sb = api.operation_sweep(b, 1, key=(4, 5), counter=c + (10 if merged else 0))
            assert np.array_equal(sa, sb) and sa[:, 0].sum() > 0
            assert _same(a, b)
    a.close()
    b.close()
