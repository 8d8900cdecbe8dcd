"""GPU parity of the replica weather features (``esm_replica_features``, src/replicas.jl:250-273)
with the oracle's restatement on the device's own weather and regimes: FP64 within 1e-12
relative, FP32 within the precision of the stored weather (1e-5 of the feature's scale), the
regime frequencies and switches exactly; and the driver's ``features`` record."""
import numpy as np
import pytest

import oracle
from esmsampling import FP32, FP64, api, synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [FP64, FP32])
@pytest.mark.parametrize("N,T,R,K", [(8, 300, 5, 2), (40, 8760, 16, 4), (6, 5, 3, 3)])
def test_replica_features_match_oracle(prec, N, T, R, K):
    print(f"replica features prec={prec} N={N} T={T} R={R} K={K}", flush=True)
    s = synth.make_system(N=N, L=max(N - 1, int(1.7 * N)), T=T, seed=4, K=K)
    dev = api.Replicas(s, R, prec)
    api.sample_weather(dev, key=(7, 2), counter=3)
    kref = s.k0 * np.random.default_rng(2).uniform(0.5, 1.5, s.k0.shape)
    f = api.replica_features(dev, kref)
    m, _ = api.weather_state(dev)
    _, _, x = dev.download()
    f_ = oracle.weather_features(s.load, x, kref, m, K)
    assert f.shape == f_.shape == (R, 3 + 5 + K)
    names = api.replica_feature_names(s)
    assert len(names) == f.shape[1]
    ireg = names.index("regime_1")
    np.testing.assert_array_equal(f[:, ireg:], f_[:, ireg:])
    if prec == FP64:
        np.testing.assert_allclose(f[:, :ireg], f_[:, :ireg], rtol=1e-12, atol=1e-12)
    else:
        # the oracle reads the device's FP32 weather exactly, so only the sums' order differs
        scale = np.abs(f_[:, :ireg]).max(axis=0) + 1e-3
        assert np.all(np.abs(f[:, :ireg] - f_[:, :ireg]) <= 1e-9 * scale)
    dev.close()


def test_features_need_weather():
    s = synth.make_system(N=5, L=6, T=24, seed=1)
    dev = api.Replicas(s, 2, FP64)
    with pytest.raises(RuntimeError, match="no weather"):
        api.replica_features(dev, s.k0)
    dev.close()
