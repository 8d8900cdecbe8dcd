"""esmsampling — MI355X-native per-sweep path of the ESM-Sampling replica sampler.

Host-side mirror of the reference's function API over the C ABI of ``libesmhip.so``
(include/esm.h). ``synth`` builds the seeded synthetic system; ``api`` holds the reference-
named operations on a batch of device-resident replicas; ``replicas`` is the replica–Langevin
driver over them (``run_replicas``) and ``diagnostics`` the convergence report and ESS of its
traces. The HIP library is loaded on first use (``_lib.load``) and raises if it is missing.
"""
from . import api, synth  # noqa: F401
from ._lib import FP32, FP64, EsmError  # noqa: F401

__all__ = ["synth", "api", "FP32", "FP64", "EsmError"]
