"""GPU parity of the cross-node storage pairs (src/storage_cross.jl) with the oracle.

`sxpair` (a family of the local sweep: one thread per step and replica) and `sxlong` (the interval
pairs of the long-range sweep: a device schedule of rounds, one wave per move and replica). FP64:
the same move statistics and states to round-off (1e-9 of the scale), as the other families;
FP32: the caches equal a rebuild to FP32 round-off. Networks of 12 nodes and more, so that the
balls of radius one around a pair leave leak nodes (the compensated kinks) as well as pairs whose
ball covers the component; hmin = 0 lets both carriers of a node take part (pairs at one node).
"""
import numpy as np
import pytest

import oracle
from esmsampling import FP32, FP64, api, synth

pytestmark = pytest.mark.gpu


def _state(s, R, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    v[0] = rng.normal(0, 200.0, (s.N, s.T, R))
    v[2:] = rng.uniform(0.05, 0.95, v[2:].shape)
    return v, x


def _check(s, vd, rd, Fd, st, v_, r_, F_, st_):
    assert np.array_equal(st[:, :4], st_[:, :4]), (st[13:], st_[13:])
    for f, scale in ((0, max(np.abs(v_[0]).max(), 1.0)), (1, 1.0), (2, 1.0), (3, 1.0)):
        np.testing.assert_allclose(vd[f], v_[f], atol=1e-9 * scale, err_msg=f"field {f}")
    np.testing.assert_allclose(rd, r_, atol=1e-9 * np.abs(s.load).max())
    np.testing.assert_allclose(Fd, F_, atol=1e-9 * max(np.abs(F_).max(), 1.0))


@pytest.mark.parametrize("R,T,hmin", [(4, 47, 24.0), (64, 24, 24.0), (3, 31, 0.0)])
def test_sxpair_fp64_matches_oracle(R, T, hmin):
    """sxpair alone (every plain weight 0: all draws are storage pairs; no atoms, no storage
    sweeps)."""
    print(f"sxpair fp64 R={R} T={T} hmin={hmin}", flush=True)
    s = synth.make_system(N=12, L=18, T=T, dt=1.0, seed=21)
    v, x = _state(s, R, seed=R + T)
    dev = api.Replicas(s, R, FP64)
    dev.upload(v=v, k=s.k0, x=x)
    api.sweep_options(dev, extra=2.0, weights=np.zeros(9), atoms=False, p_storage_ms=0.0)
    api.storage_long_options(dev, p=0.0, nmoves=0, hmin=hmin)
    api.storage_cross_options(dev, p_xpair=0.1, nxlong=0)
    st = api.operation_sweep(dev, 2, key=(9, 3), counter=4)
    vd, _, _ = dev.download()
    rd, Fd = dev.caches()
    v_, r_, F_, st_ = oracle.Oracle(s).sweeps(v, s.k0, 2, key=(9, 3), counter=4, x=x, threads=4, extra=2.0,
                                               weights=np.zeros(9), atoms=False, p_storage_ms=0.0, p_storage_long=0.0,
                                               long_moves=0, long_hours=hmin, p_xpair=0.1, xlong_moves=0)
    assert st_[13, 0] > 0 and np.all(st_[:13, 0] == 0) and st_[13, 1] > 0
    _check(s, vd, rd, Fd, st, v_, r_, F_, st_)
    dev.close()


@pytest.mark.parametrize("R,T,hmin", [(3, 47, 24.0), (64, 30, 24.0), (2, 300, 24.0), (3, 40, 0.0)])
def test_sxlong_fp64_matches_oracle(R, T, hmin):
    """The long-range sweep with its interval pairs (level moves, no plateaus, 6 pairs per unit)."""
    print(f"sxlong fp64 R={R} T={T} hmin={hmin}", flush=True)
    s = synth.make_system(N=12, L=18, T=T, dt=1.0, seed=23)
    v, x = _state(s, R, seed=2 * R + T)
    dev = api.Replicas(s, R, FP64)
    dev.upload(v=v, k=s.k0, x=x)
    api.sweep_options(dev, extra=0.0, atoms=False, p_storage_ms=0.0)
    api.storage_long_options(dev, p=1.0, nmoves=0, hmin=hmin)
    api.storage_cross_options(dev, p_xpair=0.0, nxlong=6)
    st = api.operation_sweep(dev, 1, key=(6, 2), counter=9)
    vd, _, _ = dev.download()
    rd, Fd = dev.caches()
    v_, r_, F_, st_ = oracle.Oracle(s).sweeps(v, s.k0, 1, key=(6, 2), counter=9, x=x, threads=4, extra=0.0,
                                               atoms=False, p_storage_ms=0.0, p_storage_long=1.0, long_moves=0,
                                               long_hours=hmin, p_xpair=0.0, xlong_moves=6)
    assert st_[14, 0] > 0 and st_[14, 1] > 0
    _check(s, vd, rd, Fd, st, v_, r_, F_, st_)
    dev.close()


def test_full_sweep_with_storage_pairs_fp64_matches_oracle_40_nodes():
    """Every family and both storage sweeps with the reference configuration's pairs (0.1, 8)."""
    print("full sweep with storage pairs fp64 40 nodes", flush=True)
    s = synth.make_system(N=40, L=70, T=48, dt=1.0, seed=5)
    R = 3
    v, x = _state(s, R, seed=31)
    dev = api.Replicas(s, R, FP64)
    dev.upload(v=v, k=s.k0, x=x)
    api.sweep_options(dev, p_storage_ms=1.0)
    api.storage_cross_options(dev, p_xpair=0.1, nxlong=8)
    st = api.operation_sweep(dev, 1, key=(2, 8), counter=3)
    vd, _, _ = dev.download()
    rd, Fd = dev.caches()
    v_, r_, F_, st_ = oracle.Oracle(s).sweeps(v, s.k0, 1, key=(2, 8), counter=3, x=x, threads=8, p_storage_ms=1.0,
                                               p_xpair=0.1, xlong_moves=8)
    assert st_[13, 0] > 0 and st_[14, 0] > 0
    _check(s, vd, rd, Fd, st, v_, r_, F_, st_)
    dev.close()


def test_storage_pairs_fp64_256_nodes_with_the_largest_templates():
    """At 256 nodes / 450 lines the balls around hub nodes leave up to 46 leak nodes and 66
    balances per end (the oracle's template sizes), beyond one kink slot and two balance slots of
    the 64-lane form: both pair kernels take two kink slots and three balance slots per lane, and
    sxpair and sxlong give the oracle's statistics and states (FP64)."""
    print("storage pairs fp64 256 nodes", flush=True)
    s = synth.make_system(N=256, L=450, T=12, dt=1.0, seed=1)
    o = oracle.Oracle(s)
    sizes = o.sx_sizes()
    assert sizes[:, 0].max() > 32 and sizes[:, 2].max() > 64
    R = 2
    v, x = _state(s, R, seed=5)
    dev = api.Replicas(s, R, FP64)
    dev.upload(v=v, k=s.k0, x=x)
    api.sweep_options(dev, extra=2.0, weights=np.zeros(9), atoms=False, p_storage_ms=0.0)
    api.storage_long_options(dev, p=1.0, nmoves=0, hmin=24.0)
    api.storage_cross_options(dev, p_xpair=0.1, nxlong=2)
    st = api.operation_sweep(dev, 1, key=(5, 5), counter=2)
    vd, _, _ = dev.download()
    rd, Fd = dev.caches()
    v_, r_, F_, st_ = o.sweeps(v, s.k0, 1, key=(5, 5), counter=2, x=x, threads=8, extra=2.0, weights=np.zeros(9),
                               atoms=False, p_storage_ms=0.0, p_storage_long=1.0, long_moves=0, long_hours=24.0,
                               p_xpair=0.1, xlong_moves=2)
    assert st_[13, 0] > 0 and st_[14, 0] > 0 and st[13, 3] == 0 and st[14, 3] == 0
    _check(s, vd, rd, Fd, st, v_, r_, F_, st_)
    dev.close()


def test_storage_pairs_fp32_consistent():
    """FP32 at 64 replicas: the pair families make the oracle's moves (counts exact, acceptances
    within round-off-driven differences) and the caches stay within FP32 round-off of a rebuild."""
    print("storage pairs fp32", flush=True)
    s = synth.make_system(N=12, L=18, T=48, dt=1.0, seed=21)
    R = 64
    v, x = _state(s, R, seed=7)
    x = x.astype(np.float32).astype(np.float64)
    v = v.astype(np.float32).astype(np.float64)
    dev = api.Replicas(s, R, FP32)
    dev.upload(v=v, k=s.k0, x=x)
    api.sweep_options(dev, extra=2.0, weights=np.zeros(9), atoms=False, p_storage_ms=0.0)
    api.storage_cross_options(dev, p_xpair=0.1, nxlong=8)
    st = api.operation_sweep(dev, 2, key=(4, 4), counter=1)
    _, _, drift = api.LocalCost(dev)
    v_, r_, F_, st_ = oracle.Oracle(s).sweeps(v, s.k0, 2, key=(4, 4), counter=1, x=x, threads=8, extra=2.0,
                                               weights=np.zeros(9), atoms=False, p_storage_ms=0.0, p_xpair=0.1,
                                               xlong_moves=8)
    for f in (13, 14):
        assert st_[f, 0] > 0 and st[f, 0] == st_[f, 0] and st[f, 2] == 0 and st[f, 3] == 0
        assert abs(st[f, 1] - st_[f, 1]) <= max(3, 0.02 * st_[f, 1])
    assert np.all(drift < 1e-3 * np.abs(s.load).max())
    dev.close()


def test_storage_pair_options_are_checked():
    s = synth.make_system(N=8, L=10, T=12, dt=1.0, seed=1)
    dev = api.Replicas(s, 2, FP64)
    with pytest.raises(Exception, match="p_xpair"):
        api.storage_cross_options(dev, p_xpair=-1.0, nxlong=0)
    with pytest.raises(Exception, match="nxlong"):
        api.storage_cross_options(dev, p_xpair=0.1, nxlong=300)
    api.storage_cross_options(dev, p_xpair=0.1, nxlong=8)
    dev.close()
