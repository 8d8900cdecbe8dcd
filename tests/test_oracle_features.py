"""The oracle's replica weather features (``replica_features``, src/replicas.jl:250-273) pinned by
closed forms: constant weather with one load spike gives the residual load's mean, peak and
worst 7-step mean exactly (a window of the wrong length or a non-cyclic window fails), clamping
outside [0, 1], the regime frequencies and switches of enumerated paths, and the sign and
weighting of the reference capacities."""
import numpy as np

import oracle


def _case(N=3, T=20, R=2, a=0.4, spikes=((5, 700.0),), d0=300.0):
    load = np.full((N, T), d0 / N)
    for t, s in spikes:
        load[0, t] += s
    x = np.full((3, N, T, R), a)
    kref = np.zeros((6, N))
    kref[0] = [10.0, 20.0, 30.0][:N]
    kref[1] = 5.0
    kref[2] = 1.0
    m = np.zeros((T, R), dtype=np.int32)
    return load, x, kref, m


def test_constant_weather_one_spike():
    load, x, kref, m = _case()
    T = load.shape[1]
    f = oracle.weather_features(load, x, kref, m, K=2)
    ksum = kref[:3].sum()
    base = (300.0 - 0.4 * ksum) / 1e3
    for j in range(2):
        assert np.isclose(f[j, 0], (300.0 + 700.0 / T) / 1e3, rtol=1e-14)       # mean load
        assert np.allclose(f[j, 1:4], 0.4, rtol=1e-14)                           # cf
        assert np.isclose(f[j, 4], base + 0.7 / T, rtol=1e-13)                   # mean residual
        assert np.isclose(f[j, 5], base + 0.7, rtol=1e-13)                       # peak
        assert np.isclose(f[j, 6], base + 0.7 / 7, rtol=1e-13)                   # worst 7-step mean
        assert np.array_equal(f[j, 7:], [1.0, 0.0, 0.0])                         # regimes, switches


def test_seven_step_window_is_cyclic():
    # two spikes at the last and the first step: only a window that wraps holds both
    load, x, kref, m = _case(T=30, spikes=((0, 140.0), (29, 140.0)))
    f = oracle.weather_features(load, x, kref, m, K=2)
    base = (300.0 - 0.4 * kref[:3].sum()) / 1e3
    assert np.isclose(f[0, 6], base + 0.28 / 7, rtol=1e-13)
    assert np.isclose(f[0, 5], base + 0.14, rtol=1e-13)


def test_clamped_capacity_factors_and_negative_load():
    load, x, kref, m = _case(N=2, T=8, R=1, spikes=())
    x[0] = 1.7      # solar above 1: counts as 1
    x[1] = -0.3     # onwind below 0: counts as 0
    load[1] = -50.0  # negative load enters D as 0
    f = oracle.weather_features(load, x, kref, m, K=1)
    assert np.allclose(f[0, 1:4], [1.0, 0.0, 0.4])
    d = load[0].mean()
    assert np.isclose(f[0, 0], d / 1e3)
    res = (load[0, 0] - (kref[0].sum() * 1.0 + kref[2].sum() * 0.4)) / 1e3
    assert np.isclose(f[0, 4], res) and np.isclose(f[0, 5], res) and np.isclose(f[0, 6], res)


def test_regime_frequencies_and_switches():
    load, x, kref, m = _case(T=6, R=3)
    m[:, 0] = [0, 0, 1, 1, 1, 0]
    m[:, 1] = [2, 1, 0, 1, 2, 2]
    m[:, 2] = 1
    f = oracle.weather_features(load, x, kref, m, K=3)
    assert np.allclose(f[0, 7:], [3 / 6, 3 / 6, 0.0, 2.0])
    assert np.allclose(f[1, 7:], [1 / 6, 2 / 6, 3 / 6, 4.0])
    assert np.allclose(f[2, 7:], [0.0, 1.0, 0.0, 0.0])


def test_reference_capacities_weight_each_node():
    # weather at one node only: the residual falls by kref at that node times its factor
    load, x, kref, m = _case(N=3, T=10, R=1, a=0.0, spikes=())
    x[2, 1, :, 0] = 0.5        # offwind at node 1
    f = oracle.weather_features(load, x, kref, m, K=2)
    assert np.isclose(f[0, 4], (300.0 - kref[2, 1] * 0.5) / 1e3)
    assert np.isclose(f[0, 3], 0.5 / 3)
