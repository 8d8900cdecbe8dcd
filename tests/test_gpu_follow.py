"""GPU parity of the dispatch following with the oracle: bitwise the same committed capacities,
the gas utilizations and residual caches within the precision tolerance."""
import numpy as np
import pytest

import oracle
from esmsampling import FP32, FP64, api, synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [FP64, FP32])
def test_follow_matches_oracle(prec):
    print(f"follow prec={prec}", flush=True)
    s = synth.make_system(N=10, L=16, T=48, dt=1.0, seed=4)
    R = 4
    rng = np.random.default_rng(4)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    v[1] = rng.uniform(0.05, 0.95, v[1].shape)
    v[2] = rng.uniform(0.1, 0.9, v[2].shape)
    v[3] = rng.uniform(0.1, 0.9, v[3].shape)
    if prec == FP32:
        v, x = (a.astype(np.float32).astype(np.float64) for a in (v, x))
    k = s.k0.copy()
    knew = k * np.exp(rng.normal(0, 0.03, k.shape))
    dev = api.Replicas(s, R, prec)
    dev.upload(v=v, k=k, x=x)
    kout = api.follow_capacities(dev, knew, kmin=1.0)
    vd, kd, _ = dev.download()
    rd, _ = dev.caches()
    v_, k_ = oracle.Oracle(s).follow(v, k, x, knew, kmin=1.0)
    assert np.array_equal(kout, k_) and np.array_equal(kd, k_)
    tol = 1e-10 if prec == FP64 else 2e-5
    np.testing.assert_allclose(vd, v_, atol=tol)
    r_ = oracle.Oracle(s).energy(v_, k_, x)[2]
    np.testing.assert_allclose(rd, r_, atol=(1e-9 if prec == FP64 else 1e-5) * np.abs(s.load).max())
    dev.close()
