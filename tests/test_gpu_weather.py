"""GPU parity of the weather process with the oracle: the same Philox numbers give bit-identical
regimes and move decisions; anomalies, weather and the compensated gas utilization within the
precision tolerance (FP64 1e-12, FP32 1e-5 of the quantity's scale)."""
import numpy as np
import pytest

import oracle
from esmsampling import FP32, FP64, api, synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [FP64, FP32])
@pytest.mark.parametrize("K", [2, 4])
def test_weather_init_matches_oracle(prec, K):
    print(f"weather init prec={prec} K={K}", flush=True)
    s = synth.make_system(N=8, L=12, T=500, seed=2, K=K)
    R = 6
    dev = api.Replicas(s, R, prec)
    api.sample_weather(dev, key=(11, 5), counter=7)
    m, z = api.weather_state(dev)
    _, _, x = dev.download()
    m_, z_, x_ = oracle.Oracle(s).weather_init(R, key=(11, 5), counter=7)
    assert np.array_equal(m, m_)
    tol = 1e-12 if prec == FP64 else 2e-5
    np.testing.assert_allclose(z, z_, atol=tol * 10)
    np.testing.assert_allclose(x, x_, atol=tol)
    dev.close()


@pytest.mark.parametrize("prec", [FP64, FP32])
@pytest.mark.parametrize("compensate,K", [(True, 2), (False, 2), (True, 4)])
def test_weather_step_matches_oracle(prec, compensate, K):
    print(f"weather step prec={prec} compensate={compensate} K={K}", flush=True)
    s = synth.make_system(N=8, L=12, T=700, seed=3, K=K)
    R = 8
    o = oracle.Oracle(s)
    m, z, x = o.weather_init(R, key=(2, 2), counter=0)
    if prec == FP32:
        z = z.astype(np.float32).astype(np.float64)
    x = synth.weather_from_z(s, z)
    if prec == FP32:
        x = x.astype(np.float32).astype(np.float64)
    v = synth.heuristic_state(s, x)
    if prec == FP32:
        v = v.astype(np.float32).astype(np.float64)
    k = s.k0
    dev = api.Replicas(s, R, prec)
    dev.upload(v=v, k=k, x=x)
    api.set_weather(dev, m, z)
    dev.upload(x=x)           # the uploaded x (rounded like the oracle's input)
    theta = np.array([0.3, 0.0, 0.3, 0.0, 0.1, 0.5, 0.0, 0.2])
    wp = np.array([0.0, 1.0, 1.0, 0.0, 0.5, 0.5, 0.5, 1.0])
    for it in range(3):
        moved = api.weather_process_step(dev, theta, wp, window=60, compensate=compensate, key=(5, 6), counter=it)
        m, z, x, v, moved_ = o.weather_step(m, z, x, v, k, theta, wp, window=60, compensate=compensate,
                                            key=(5, 6), counter=it)
        assert np.array_equal(moved, moved_)
    md, zd = api.weather_state(dev)
    vd, _, xd = dev.download()
    assert np.array_equal(md, m)
    assert np.all(moved_[3] == 0)
    tol = 1e-11 if prec == FP64 else 5e-5
    np.testing.assert_allclose(zd, z, atol=tol * 10)
    np.testing.assert_allclose(xd, x, atol=tol)
    np.testing.assert_allclose(vd, v, atol=tol, rtol=tol)
    # the residual cache follows the weather and the re-dispatch
    r, F = dev.caches()
    r_ = o.energy(v, k, x)[2]
    np.testing.assert_allclose(r, r_, atol=(1e-9 if prec == FP64 else 1e-5) * np.abs(s.load).max())
    dev.close()


@pytest.mark.parametrize("prec,K", [(FP64, 2), (FP32, 2), (FP64, 4)])
def test_weather_ffbs_matches_oracle(prec, K):
    """The centred FFBS of the regimes: from the same anomalies and Philox numbers the regime paths
    are those of the oracle's ffbs (src/weather.jl:160-174), replica for replica; masked-out
    replicas keep their paths; the anomalies and the weather do not change."""
    print(f"weather ffbs prec={prec} K={K}", flush=True)
    s = synth.make_system(N=8, L=12, T=400, seed=5, K=K)
    R = 6
    dev = api.Replicas(s, R, prec)
    api.sample_weather(dev, key=(4, 4), counter=1)
    m0, z0 = api.weather_state(dev)
    m_start = m0.copy()
    _, _, x0 = dev.download()
    mask = np.array([1, 0, 1, 1, 0, 1], dtype=np.int32)
    for c in range(3):
        api.weather_ffbs(dev, mask=mask, key=(4, 5), counter=c)
        m1, z1 = api.weather_state(dev)
        m_, _ = oracle.Oracle(s).weather_ffbs(m0, z0, mask=mask, key=(4, 5), counter=c)
        assert np.array_equal(m1, m_)
        assert np.array_equal(m1[:, mask == 0], m0[:, mask == 0])
        assert np.array_equal(z1, z0)
        m0 = m1
    assert (m1[:, mask == 1] != m_start[:, mask == 1]).any()
    _, _, x1 = dev.download()
    assert np.array_equal(x1, x0)
    dev.close()
