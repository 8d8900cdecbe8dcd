import os
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "esm-sampling_amd"))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (MI355X) and the built HIP library")


def _has_gpu() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _has_gpu():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def small_system():
    from esmsampling import synth
    return synth.make_system(N=6, L=8, T=48, dt=1.0, seed=3)


@pytest.fixture(scope="session")
def sys40():
    from esmsampling import synth
    return synth.make_system(N=40, L=70, T=240, dt=1.0, seed=1)


def random_state(sys, R, seed=0, spread=True):
    """A random state exercising every branch: angles that overload some lines, utilizations and
    states of charge partly outside [0, 1], raw capacity factors partly outside [0, 1]."""
    import numpy as np
    from esmsampling import synth
    rng = np.random.default_rng(seed)
    N, T = sys.N, sys.T
    v = np.empty((synth.NV, N, T, R))
    v[0] = rng.normal(0.0, 800.0, (N, T, R))
    v[1] = rng.uniform(-0.1, 1.1, (N, T, R)) if spread else rng.uniform(0.05, 0.95, (N, T, R))
    v[2] = rng.uniform(-0.05, 1.05, (N, T, R)) if spread else rng.uniform(0.1, 0.9, (N, T, R))
    v[3] = rng.uniform(-0.05, 1.05, (N, T, R)) if spread else rng.uniform(0.1, 0.9, (N, T, R))
    x = rng.uniform(-0.05, 1.05, (synth.NREN, N, T, R)) if spread else rng.uniform(0.01, 0.99, (synth.NREN, N, T, R))
    k = sys.k0 * rng.uniform(0.5, 1.5, sys.k0.shape)
    if spread:
        k[0, 0] = sys.khi[0, 0] * 1.2      # above the box: clamped, penalised, masked score
    return v, k, x
