"""GPU parity of the energy, flow and residual caches, scores and Langevin step against the oracle.

Tolerances: FP64 state 1e-12 relative, FP32 state 1e-5 relative; the
oracle evaluates the same (FP32-rounded, in FP32 mode) state in FP64.
"""
import sys

import numpy as np
import pytest

import oracle
from conftest import random_state
from esmsampling import FP32, FP64, synth
from esmsampling import api

pytestmark = pytest.mark.gpu


def _round32(*arrs):
    return [a.astype(np.float32).astype(np.float64) for a in arrs]


@pytest.mark.parametrize("prec", [FP64, FP32])
@pytest.mark.parametrize("spread", [True, False])
def test_energy_parts_and_caches(sys40, prec, spread):
    print(f"energy parity prec={prec} spread={spread}", flush=True)
    s = sys40
    R = 5
    v, k, x = random_state(s, R, seed=7, spread=spread)
    if prec == FP32:
        v, x = _round32(v, x)
    dev = api.Replicas(s, R, prec)
    dev.upload(v=v, k=k, x=x)
    parts, ph = api.energy_cost(dev, per_hour=True)
    r, F = dev.caches()
    op_, ph_, r_, F_ = oracle.Oracle(s).energy(v, k, x)
    rtol = 1e-12 if prec == FP64 else 1e-5
    scale_r = np.abs(s.load).max()
    np.testing.assert_allclose(F, F_, rtol=rtol, atol=rtol * np.abs(F_).max())
    np.testing.assert_allclose(r, r_, rtol=rtol, atol=rtol * scale_r)
    for p in (0, 1, 2, 3, 6, 7):
        np.testing.assert_allclose(parts[:, p], op_[:, p], rtol=rtol, err_msg=api.PART_NAMES[p])
    # shed/dump MWh: relative to the total residual volume (their terms cancel nothing)
    vol = op_[:, 4] + op_[:, 5]
    np.testing.assert_allclose(parts[:, 4], op_[:, 4], atol=rtol * vol.max())
    np.testing.assert_allclose(parts[:, 5], op_[:, 5], atol=rtol * vol.max())
    np.testing.assert_allclose(ph, ph_, rtol=rtol, atol=rtol * np.abs(ph_).max())
    dev.close()


@pytest.mark.parametrize("prec", [FP64, FP32])
def test_rebuild_drift_zero(sys40, prec):
    s = sys40
    R = 3
    v, k, x = random_state(s, R, seed=8)
    dev = api.Replicas(s, R, prec)
    dev.upload(v=v, k=k, x=x)
    r0, F0, drift = api.LocalCost(dev)
    assert np.all(drift == 0.0)     # the upload rebuilt the caches: the rebuild is bitwise the same
    dev.close()


@pytest.mark.parametrize("prec", [FP64, FP32])
@pytest.mark.parametrize("measure", ["utilization", "physical"])
def test_scores(sys40, prec, measure):
    print(f"scores parity prec={prec} measure={measure}", flush=True)
    s = sys40
    R = 4
    v, k, x = random_state(s, R, seed=9)
    if prec == FP32:
        v, x = _round32(v, x)
    dev = api.Replicas(s, R, prec)
    dev.upload(v=v, k=k, x=x)
    api.score_measure(dev, measure)
    sk, fk = api.capacity_scores(dev)
    sk_, fk_, sa = oracle.Oracle(s).scores(v, k, x, measure=measure)
    np.testing.assert_allclose(fk, fk_, rtol=1e-14)
    tol = 1e-12 if prec == FP64 else 1e-5
    # each entry is a sum of terms of both signs: the round-off scale is the sum of their magnitudes
    assert np.all(np.abs(sk - sk_) <= tol * sa + 1e-300)
    dev.close()


def test_langevin_matches_oracle(sys40):
    s = sys40
    R = 4
    v, k, x = random_state(s, R, seed=10)
    dev = api.Replicas(s, R, FP64)
    dev.upload(v=v, k=k, x=x)
    sk, fk = api.capacity_scores(dev)
    rng = np.random.default_rng(3)
    w = api.uniform_weights([True, False, True, True])
    xi = rng.standard_normal((6, s.N))
    eta = np.array([5e-4, 5e-4, 5e-4, 5e-4, 5e-4, 0.0])
    ks = np.array([1000.0, 500.0, 200.0, 300.0, 100.0, 100.0])
    knew, nc = api.langevin_capacities(dev, w, xi, eta, ks, kmin=1.0, clip=0.02, box=True)
    knew_, nc_ = oracle.Oracle(s).langevin(k, sk, fk, w, xi, eta, ks, kmin=1.0, clip=0.02, box=True)
    np.testing.assert_allclose(knew, knew_, rtol=1e-13, atol=1e-9)
    assert nc == nc_
    assert np.all(knew[5] == k[5])      # eta = 0: frozen
    dev.close()
