"""The allocation-free Langevin step (esm_langevin_noalloc) through the C ABI: the same new
capacities and clip count as esm_langevin, bit for bit, on two calls in a row on one device —
the second with a smaller replica set and network, whose clip count would show a counter left
over from the first call — and a clear error for a set beyond its fixed scratch."""
import ctypes as C

import numpy as np
import pytest

from conftest import random_state
from esmsampling import FP32, FP64, api, synth
from esmsampling._lib import dptr, f64

pytestmark = pytest.mark.gpu

ETA = f64(np.full(6, 5e-4))
KS = f64([1000.0, 500.0, 200.0, 300.0, 100.0, 100.0])


def _step(dev, name, w, xi, clip):
    knew = np.empty((6, dev.N))
    nc = C.c_int32(-1)
    rc = getattr(dev.lib, name)(dev.handle, dptr(w), dptr(xi), dptr(ETA), dptr(KS), 1.0, clip, 1, dptr(knew),
                                C.byref(nc))
    assert rc == 0, dev.lib.esm_last_error()
    return knew, nc.value


def test_noalloc_langevin_equals_langevin_on_two_calls():
    clips = []
    for N, L, T, R, clip, prec, seed in ((40, 70, 24, 64, 1e-4, FP32, 1), (10, 16, 12, 8, 0.02, FP64, 2)):
        print(f"langevin noalloc N={N} R={R} clip={clip}", flush=True)
        s = synth.make_system(N=N, L=L, T=T, dt=1.0, seed=seed)
        v, k, x = random_state(s, R, seed=seed)
        dev = api.Replicas(s, R, prec)
        dev.upload(v=v, k=k, x=x)
        api.capacity_scores(dev)
        rng = np.random.default_rng(seed)
        w = f64(api.uniform_weights(rng.random(R) < 0.7))
        xi = f64(rng.standard_normal((6, N)))
        a = _step(dev, "esm_langevin", w, xi, clip)
        b = _step(dev, "esm_langevin_noalloc", w, xi, clip)
        assert np.array_equal(a[0], b[0]) and a[1] == b[1], (a[1], b[1])
        # and the allocating form is unaffected by the shared scratch's contents
        c = _step(dev, "esm_langevin", w, xi, clip)
        assert np.array_equal(a[0], c[0]) and a[1] == c[1]
        clips.append(b[1])
        dev.close()
    assert clips[0] > clips[1] >= 0          # the first call's count did not carry over


def test_noalloc_langevin_rejects_sets_beyond_its_scratch():
    print("langevin noalloc size limit", flush=True)
    s = synth.make_system(N=256, L=450, T=2, dt=1.0, seed=3)
    R = 5200                                  # R + 12 N + 13 = 8285 > 8192
    dev = api.Replicas(s, R, FP32)
    w = f64(np.full(R, 1.0 / R))
    xi = f64(np.zeros((6, s.N)))
    knew = np.empty((6, s.N))
    rc = dev.lib.esm_langevin_noalloc(dev.handle, dptr(w), dptr(xi), dptr(ETA), dptr(KS), 1.0, 0.02, 1, dptr(knew),
                                      None)
    assert rc == -1 and b"scratch" in dev.lib.esm_last_error()
    dev.close()
