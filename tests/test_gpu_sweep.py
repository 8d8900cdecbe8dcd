"""GPU parity of the exact line-move sweep with the oracle.

FP64: with the same Philox numbers every (site, replica) takes the same decision; after whole
sweeps the state and the caches agree to round-off (a different rounding in the last bits can
move a draw by that much, never change a decision except at a CDF boundary: tolerance 1e-9 of
the scale), and the move statistics agree exactly.
FP32: the same statistics within round-off-driven differences (distributional); the caches equal
a rebuild to FP32 round-off.
"""
import numpy as np
import pytest

import oracle
from conftest import random_state
from esmsampling import FP32, FP64, api, synth

pytestmark = pytest.mark.gpu


def _sys(N=10, L=16, T=48, seed=8):
    return synth.make_system(N=N, L=L, T=T, dt=1.0, seed=seed)


def test_schedule_matches_oracle():
    s = _sys()
    dev = api.Replicas(s, 2, FP64)
    col, nc = api.schedule(dev)
    col_, nc_ = oracle.Oracle(s).colouring()
    assert nc == nc_ and np.array_equal(col, col_)
    dev.close()


@pytest.mark.parametrize("nsweeps", [1, 2])
def test_sweep_fp64_matches_oracle(nsweeps):
    print(f"sweep fp64 nsweeps={nsweeps}", flush=True)
    s = _sys()
    R = 5
    v, k, x = random_state(s, R, seed=12, spread=False)
    dev = api.Replicas(s, R, FP64)
    dev.upload(v=v, k=k, x=x)
    st = api.operation_sweep(dev, nsweeps, key=(3, 4), counter=10)
    vd, _, _ = dev.download()
    rd, Fd = dev.caches()
    v_, r_, F_, st_ = oracle.Oracle(s).sweeps(v, k, nsweeps, key=(3, 4), counter=10, x=x, threads=4)
    assert np.array_equal(st[:, :4], st_[:, :4])
    for f, scale in ((0, np.abs(v_[0]).max()), (1, 1.0), (2, 1.0), (3, 1.0)):
        np.testing.assert_allclose(vd[f], v_[f], atol=1e-9 * scale, err_msg=f"field {f}")
    np.testing.assert_allclose(rd, r_, atol=1e-9 * np.abs(s.load).max())
    np.testing.assert_allclose(Fd, F_, atol=1e-9 * np.abs(F_).max())
    dev.close()


def test_sweep_fp64_larger_network():
    print("sweep fp64 40 nodes", flush=True)
    s = synth.make_system(N=40, L=70, T=96, dt=1.0, seed=1)
    R = 3
    rng = np.random.default_rng(2)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    v[0] = rng.normal(0, 300.0, (s.N, s.T, R))
    k = s.k0
    dev = api.Replicas(s, R, FP64)
    dev.upload(v=v, k=k, x=x)
    st = api.operation_sweep(dev, 1, key=(9, 9), counter=0)
    vd, _, _ = dev.download()
    v_, r_, F_, st_ = oracle.Oracle(s).sweeps(v, k, 1, key=(9, 9), counter=0, x=x, threads=4)
    assert np.array_equal(st[:, :4], st_[:, :4])
    np.testing.assert_allclose(vd[1:], v_[1:], atol=1e-9)
    np.testing.assert_allclose(vd[0], v_[0], atol=1e-9 * np.abs(v_[0]).max())
    dev.close()


def test_sweep_fp32_consistent():
    print("sweep fp32", flush=True)
    s = _sys()
    R = 16
    v, k, x = random_state(s, R, seed=13, spread=False)
    v, x = (a.astype(np.float32).astype(np.float64) for a in (v, x))
    dev = api.Replicas(s, R, FP32)
    dev.upload(v=v, k=k, x=x)
    st = api.operation_sweep(dev, 2, key=(5, 5), counter=0)
    _, _, drift = api.LocalCost(dev)
    v_, r_, F_, st_ = oracle.Oracle(s).sweeps(v, k, 2, key=(5, 5), counter=0, x=x, threads=4)
    assert np.all(st[:, 0] == st_[:, 0])
    assert np.all(st[:, 2:4] == 0)
    assert np.all(np.abs(st[:, 1] - st_[:, 1]) <= np.maximum(3, 0.02 * st_[:, 1]))
    assert np.all(drift < 1e-3 * np.abs(s.load).max())
    dev.close()


WEIGHT_NAMES = ("unit", "pair", "xpair", "xlocc", "xloc", "potential", "cellpair", "cell", "spair")


@pytest.mark.parametrize("family", ["unit", "potential", "spair", "cell", "xloc", "xlocc", "xpair"])
def test_random_family_fp64_matches_oracle(family):
    """One random family alone (no atoms): identical move statistics, states to 1e-9."""
    print(f"random family {family} fp64", flush=True)
    s = _sys(N=10, L=16, T=47, seed=4)             # odd T: three parity classes
    R = 4
    v, k, x = random_state(s, R, seed=15, spread=False)
    w = np.zeros(9)
    w[WEIGHT_NAMES.index(family)] = 1.0
    w[WEIGHT_NAMES.index("pair")] = 0.5           # void draws take their share
    dev = api.Replicas(s, R, FP64)
    dev.upload(v=v, k=k, x=x)
    api.sweep_options(dev, extra=2.0, weights=w, atoms=False, p_storage_ms=0.0)
    st = api.operation_sweep(dev, 2, key=(8, 1), counter=5)
    vd, _, _ = dev.download()
    rd, Fd = dev.caches()
    v_, r_, F_, st_ = oracle.Oracle(s).sweeps(v, k, 2, key=(8, 1), counter=5, x=x, threads=4, extra=2.0, weights=w,
                                               atoms=False, p_storage_ms=0.0)
    assert st_[:, 0].sum() > 0
    assert np.array_equal(st[:, :4], st_[:, :4]), (st[:, :4], st_[:, :4])
    for f, scale in ((0, max(np.abs(v_[0]).max(), 1.0)), (1, 1.0), (2, 1.0), (3, 1.0)):
        np.testing.assert_allclose(vd[f], v_[f], atol=1e-9 * scale, err_msg=f"field {f}")
    np.testing.assert_allclose(rd, r_, atol=1e-9 * np.abs(s.load).max())
    np.testing.assert_allclose(Fd, F_, atol=1e-9 * max(np.abs(F_).max(), 1.0))
    dev.close()


def test_full_sweep_fp64_matches_oracle_40_nodes():
    """Atoms and every random family at the reference's rates on the 40-node network."""
    print("full sweep fp64 40 nodes", flush=True)
    s = synth.make_system(N=40, L=70, T=24, dt=1.0, seed=1)
    R = 3
    rng = np.random.default_rng(5)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    v[0] = rng.normal(0, 300.0, (s.N, s.T, R))
    dev = api.Replicas(s, R, FP64)
    dev.upload(v=v, k=s.k0, x=x)
    st = api.operation_sweep(dev, 1, key=(2, 3), counter=1)
    vd, _, _ = dev.download()
    v_, r_, F_, st_ = oracle.Oracle(s).sweeps(v, s.k0, 1, key=(2, 3), counter=1, x=x, threads=4)
    assert np.all(st_[:10, 0] > 0)              # the storage sweep runs with probability 1/2
    assert np.array_equal(st[:, :4], st_[:, :4])
    np.testing.assert_allclose(vd[1:], v_[1:], atol=1e-9)
    np.testing.assert_allclose(vd[0], v_[0], atol=1e-9 * np.abs(v_[0]).max())
    dev.close()


def test_sweep_fp64_matches_oracle_64_replicas():
    """64 replicas: every wave holds one step's replicas (the wave-uniform kernels)."""
    print("full sweep fp64 64 replicas", flush=True)
    s = synth.make_system(N=10, L=16, T=10, dt=1.0, seed=6)
    R = 64
    rng = np.random.default_rng(8)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    v[0] = rng.normal(0, 300.0, (s.N, s.T, R))
    dev = api.Replicas(s, R, FP64)
    dev.upload(v=v, k=s.k0, x=x)
    st = api.operation_sweep(dev, 2, key=(5, 6), counter=3)
    vd, _, _ = dev.download()
    rd, Fd = dev.caches()
    v_, r_, F_, st_ = oracle.Oracle(s).sweeps(v, s.k0, 2, key=(5, 6), counter=3, x=x, threads=8)
    assert np.all(st_[:10, 0] > 0)
    assert np.array_equal(st[:, :4], st_[:, :4])
    np.testing.assert_allclose(vd[1:], v_[1:], atol=1e-9)
    np.testing.assert_allclose(vd[0], v_[0], atol=1e-9 * np.abs(v_[0]).max())
    np.testing.assert_allclose(rd, r_, atol=1e-9 * np.abs(s.load).max())
    dev.close()


@pytest.mark.parametrize("prec,R", [(FP64, 3), (FP32, 64)])
def test_merged_random_launch_equals_one_launch_per_family(prec, R):
    """The table-driven families of a parity class in one launch (the default) leave states, caches
    and statistics bitwise equal to one launch per family (steps of a class share no values)."""
    print(f"merged random launch prec={prec} R={R}", flush=True)
    s = synth.make_system(N=40, L=70, T=48, dt=1.0, seed=2)
    rng = np.random.default_rng(9)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    v[0] = rng.normal(0, 300.0, (s.N, s.T, R))
    out = []
    for merged in (True, False):
        dev = api.Replicas(s, R, prec)
        dev.upload(v=v, k=s.k0, x=x)
        api.sweep_launch(dev, merged)
        st = api.operation_sweep(dev, 2, key=(4, 4), counter=5)
        out.append((st, *dev.download(), *dev.caches()))
        dev.close()
    (sa, va, _, _, ra, Fa), (sb, vb, _, _, rb, Fb) = out
    assert np.all(sa[3:10, 0] > 0)
    assert np.array_equal(sa, sb)
    assert np.array_equal(va, vb) and np.array_equal(ra, rb) and np.array_equal(Fa, Fb)


@pytest.mark.parametrize("waves", [2, 4])
def test_sweep_fp64_waves_per_step_matches_oracle(waves):
    """The merged table-driven launch with several waves per (step, 64 replicas), each taking every
    waves-th template entry, against the oracle (FP64): the same move statistics, states to 1e-9
    (the terms of ΔC are summed in another order)."""
    print(f"sweep fp64 waves per step {waves}", flush=True)
    s = synth.make_system(N=40, L=70, T=8, dt=1.0, seed=3)
    R = 64
    rng = np.random.default_rng(11)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    v[0] = rng.normal(0, 300.0, (s.N, s.T, R))
    dev = api.Replicas(s, R, FP64)
    dev.upload(v=v, k=s.k0, x=x)
    api.sweep_launch(dev, waves)
    st = api.operation_sweep(dev, 1, key=(7, 3), counter=2)
    vd, _, _ = dev.download()
    rd, Fd = dev.caches()
    v_, r_, F_, st_ = oracle.Oracle(s).sweeps(v, s.k0, 1, key=(7, 3), counter=2, x=x, threads=8)
    assert np.all(st_[6:10, 0] > 0)
    assert np.array_equal(st[:, :4], st_[:, :4])
    np.testing.assert_allclose(vd[1:], v_[1:], atol=1e-9)
    np.testing.assert_allclose(vd[0], v_[0], atol=1e-9 * np.abs(v_[0]).max())
    np.testing.assert_allclose(rd, r_, atol=1e-9 * np.abs(s.load).max())
    np.testing.assert_allclose(Fd, F_, atol=1e-9 * np.abs(F_).max())
    with pytest.raises(Exception, match="multiple of 64"):
        small = api.Replicas(s, 3, FP64)
        try:
            api.sweep_launch(small, 4)
        finally:
            small.close()
    dev.close()


@pytest.mark.parametrize("R,T,N,C", [(4, 47, 10, None), (64, 20, 10, None), (3, 300, 10, None), (8, 200, 40, 2)])
def test_storage_ms_fp64_matches_oracle(R, T, N, C):
    """The storage multiscale sweep alone (every block length below T/2, segments of 4 … 64
    lanes); with two countries of 20 nodes, the groups move 20 units at once."""
    print(f"storage_ms fp64 R={R} T={T} N={N} countries={C}", flush=True)
    s = synth.make_system(N=N, L={10: 16, 40: 70}[N], T=T, dt=1.0, seed=12, ncountry=C)
    rng = np.random.default_rng(R + T)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    v[2:] = rng.uniform(0.1, 0.9, v[2:].shape)
    dev = api.Replicas(s, R, FP64)
    dev.upload(v=v, k=s.k0, x=x)
    api.sweep_options(dev, extra=0.0, atoms=False, p_storage_ms=1.0)
    st = api.operation_sweep(dev, 1, key=(4, 9), counter=2)
    vd, _, _ = dev.download()
    rd, Fd = dev.caches()
    v_, r_, F_, st_ = oracle.Oracle(s).sweeps(v, s.k0, 1, key=(4, 9), counter=2, x=x, threads=4, extra=0.0,
                                               atoms=False, p_storage_ms=1.0)
    assert st_[10, 0] > 0 and np.all(st_[:10, 0] == 0)
    assert np.array_equal(st[:, :4], st_[:, :4]), (st[10], st_[10])
    np.testing.assert_allclose(vd[1:], v_[1:], atol=1e-9)
    np.testing.assert_allclose(rd, r_, atol=1e-9 * np.abs(s.load).max())
    dev.close()


def test_storage_ms_refuses_groups_beyond_the_lds():
    """FP64 groups of 20 units at block length 128 need 10·129·20·8 B of LDS planes, more than a
    workgroup's 160 KB: the host refuses the launch with an error naming the need (no kernel of
    that length runs); FP32 takes the same groups."""
    print("storage_ms fp64 LDS refusal", flush=True)
    # T = 400: every offset of L = 128 leaves at least two blocks, so that length always runs
    s = synth.make_system(N=40, L=70, T=400, dt=1.0, seed=12, ncountry=2)
    rng = np.random.default_rng(5)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, 2))
    v = synth.heuristic_state(s, x)
    dev = api.Replicas(s, 2, FP64)
    try:
        dev.upload(v=v, k=s.k0, x=x)
        api.sweep_options(dev, extra=0.0, atoms=False, p_storage_ms=1.0)
        with pytest.raises(Exception, match="LDS"):
            api.operation_sweep(dev, 1, key=(4, 9), counter=2)
    finally:
        dev.close()


@pytest.mark.parametrize("N,C", [(10, None), (40, 2)])
def test_storage_ms_fp32_consistent(N, C):
    print(f"storage_ms fp32 N={N} countries={C}", flush=True)
    s = synth.make_system(N=N, L={10: 16, 40: 70}[N], T=300, dt=1.0, seed=13, ncountry=C)
    R = 64
    rng = np.random.default_rng(3)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R)).astype(np.float32).astype(np.float64)
    v = synth.heuristic_state(s, x)
    v[2:] = rng.uniform(0.1, 0.9, v[2:].shape)
    v = v.astype(np.float32).astype(np.float64)
    dev = api.Replicas(s, R, FP32)
    dev.upload(v=v, k=s.k0, x=x)
    api.sweep_options(dev, extra=0.0, atoms=False, p_storage_ms=1.0)
    st = api.operation_sweep(dev, 2, key=(6, 6), counter=0)
    _, _, drift = api.LocalCost(dev)
    v_, r_, F_, st_ = oracle.Oracle(s).sweeps(v, s.k0, 2, key=(6, 6), counter=0, x=x, threads=8, extra=0.0,
                                               atoms=False, p_storage_ms=1.0)
    assert st[10, 0] == st_[10, 0] and st[10, 2] == 0 and st[10, 3] == 0
    assert abs(st[10, 1] - st_[10, 1]) <= max(3, 0.02 * st_[10, 1])
    assert np.all(drift < 1e-3 * np.abs(s.load).max())
    dev.close()


@pytest.mark.parametrize("R,T,hmin", [(3, 47, 24.0), (64, 30, 24.0), (2, 400, 24.0), (3, 47, 0.0)])
def test_storage_long_fp64_matches_oracle(R, T, hmin):
    """The long-range storage sweep alone: level moves and plateaus of every unit with at least
    hmin hours (hmin = 0: the batteries too, two units per node in one shuffled list)."""
    print(f"storage_long fp64 R={R} T={T} hmin={hmin}", flush=True)
    s = synth.make_system(N=10, L=16, T=T, dt=1.0, seed=14)
    rng = np.random.default_rng(R + T)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    v[2:] = rng.uniform(0.1, 0.9, v[2:].shape)
    dev = api.Replicas(s, R, FP64)
    dev.upload(v=v, k=s.k0, x=x)
    api.sweep_options(dev, extra=0.0, atoms=False, p_storage_ms=0.0)
    api.storage_long_options(dev, p=1.0, nmoves=6, hmin=hmin)
    st = api.operation_sweep(dev, 1, key=(5, 2), counter=7)
    vd, _, _ = dev.download()
    rd, Fd = dev.caches()
    v_, r_, F_, st_ = oracle.Oracle(s).sweeps(v, s.k0, 1, key=(5, 2), counter=7, x=x, threads=4, extra=0.0,
                                               atoms=False, p_storage_ms=0.0, p_storage_long=1.0, long_moves=6,
                                               long_hours=hmin)
    assert st_[11, 0] > 0 and st_[12, 0] > 0 and np.all(st_[:11, 0] == 0)
    assert np.array_equal(st[:, :4], st_[:, :4]), (st[11:], st_[11:])
    np.testing.assert_allclose(vd[1:], v_[1:], atol=1e-9)
    np.testing.assert_allclose(rd, r_, atol=1e-9 * np.abs(s.load).max())
    dev.close()


def test_storage_long_fp64_trajectory_in_global_memory():
    """A horizon whose trajectory does not fit the LDS (FP64, 20500 steps): the kernel reads it
    from global memory."""
    print("storage_long fp64 T=20500", flush=True)
    s = synth.make_system(N=4, L=4, T=20500, dt=1.0, seed=15)
    R = 2
    rng = np.random.default_rng(4)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    v[2:] = rng.uniform(0.1, 0.9, v[2:].shape)
    dev = api.Replicas(s, R, FP64)
    dev.upload(v=v, k=s.k0, x=x)
    api.sweep_options(dev, extra=0.0, atoms=False, p_storage_ms=0.0)
    api.storage_long_options(dev, p=1.0, nmoves=2, hmin=24.0)
    st = api.operation_sweep(dev, 1, key=(5, 3), counter=1)
    vd, _, _ = dev.download()
    v_, r_, F_, st_ = oracle.Oracle(s).sweeps(v, s.k0, 1, key=(5, 3), counter=1, x=x, threads=2, extra=0.0,
                                               atoms=False, p_storage_ms=0.0, p_storage_long=1.0, long_moves=2)
    assert st_[11, 0] == 2 * s.N * R and st_[12, 0] == s.N * R
    assert np.array_equal(st[:, :4], st_[:, :4]), (st[11:], st_[11:])
    np.testing.assert_allclose(vd[1:], v_[1:], atol=1e-9)
    dev.close()


def test_storage_long_fp32_consistent():
    print("storage_long fp32", flush=True)
    s = synth.make_system(N=10, L=16, T=300, dt=1.0, seed=16)
    R = 64
    rng = np.random.default_rng(5)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R)).astype(np.float32).astype(np.float64)
    v = synth.heuristic_state(s, x)
    v[2:] = rng.uniform(0.1, 0.9, v[2:].shape)
    v = v.astype(np.float32).astype(np.float64)
    dev = api.Replicas(s, R, FP32)
    dev.upload(v=v, k=s.k0, x=x)
    api.sweep_options(dev, extra=0.0, atoms=False, p_storage_ms=0.0)
    api.storage_long_options(dev, p=1.0, nmoves=8)
    st = api.operation_sweep(dev, 2, key=(6, 7), counter=0)
    _, _, drift = api.LocalCost(dev)
    v_, r_, F_, st_ = oracle.Oracle(s).sweeps(v, s.k0, 2, key=(6, 7), counter=0, x=x, threads=8, extra=0.0,
                                               atoms=False, p_storage_ms=0.0, p_storage_long=1.0, long_moves=8)
    for f in (11, 12):
        assert st[f, 0] == st_[f, 0] and st[f, 2] == 0 and st[f, 3] == 0
        assert abs(st[f, 1] - st_[f, 1]) <= max(3, 0.02 * st_[f, 1])
    assert np.all(drift < 1e-3 * np.abs(s.load).max())
    dev.close()
