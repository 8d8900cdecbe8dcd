"""ctypes binding of libesmhip.so (include/esm.h).

The library is loaded from this directory only; there is no fallback: every operation of the
package runs through the HIP library, and importing a function that needs it raises when the
library is missing or was built without a GPU-capable runtime.
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libesmhip.so"

NREN, NSTOR, NK, NV, NPARTS = 3, 2, 6, 4, 8
FP64, FP32 = 64, 32

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_u32p = C.POINTER(C.c_uint32)


class EsmModel(C.Structure):
    """``esm_model`` of include/esm.h (field order must match the header)."""

    _fields_ = [
        ("N", C.c_int32), ("L", C.c_int32), ("T", C.c_int32), ("dt", C.c_double),
        ("line_bus0", _ip), ("line_bus1", _ip), ("line_b", _dp), ("line_s", _dp),
        ("load", _dp), ("mc", _dp), ("ef", _dp), ("capex", _dp), ("klo", _dp), ("khi", _dp),
        ("stor_hours", _dp), ("stor_eta_in", _dp), ("stor_eta_out", _dp), ("stor_qt", _dp),
        ("c_shed", C.c_double), ("c_dump", C.c_double), ("c_viol", C.c_double), ("gamma_k", C.c_double),
        ("kappa0", C.c_double), ("beta_v", C.c_double), ("beta_k", C.c_double),
        ("K", C.c_int32), ("Ptr", _dp), ("pi0", _dp), ("phi", C.c_double), ("sigma", C.c_double),
        ("sigma0", C.c_double), ("mu", _dp), ("xbar", _dp), ("ncountry", C.c_int32), ("country", _ip),
    ]


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_ip)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def make_cmodel(sys):
    """An ``esm_model`` pointing into contiguous copies of the system's arrays (kept alive in ``keep``)."""
    keep = {
        "bus0": np.ascontiguousarray(sys.line_bus0, dtype=np.int32),
        "bus1": np.ascontiguousarray(sys.line_bus1, dtype=np.int32),
        "b": f64(sys.line_b), "s": f64(sys.line_s), "load": f64(sys.load), "mc": f64(sys.mc), "ef": f64(sys.ef),
        "capex": f64(sys.capex), "klo": f64(sys.klo), "khi": f64(sys.khi), "hours": f64(sys.stor_hours),
        "ein": f64(sys.stor_eta_in), "eout": f64(sys.stor_eta_out), "qt": f64(sys.stor_qt),
        "Ptr": f64(sys.Ptr), "pi0": f64(sys.pi0), "mu": f64(sys.mu), "xbar": f64(sys.xbar),
        "country": np.ascontiguousarray(sys.country if getattr(sys, "country", None) is not None
                                        else np.zeros(sys.N), dtype=np.int32),
    }
    m = EsmModel(
        N=sys.N, L=sys.L, T=sys.T, dt=sys.dt,
        line_bus0=iptr(keep["bus0"]), line_bus1=iptr(keep["bus1"]), line_b=dptr(keep["b"]), line_s=dptr(keep["s"]),
        load=dptr(keep["load"]), mc=dptr(keep["mc"]), ef=dptr(keep["ef"]), capex=dptr(keep["capex"]),
        klo=dptr(keep["klo"]), khi=dptr(keep["khi"]), stor_hours=dptr(keep["hours"]),
        stor_eta_in=dptr(keep["ein"]), stor_eta_out=dptr(keep["eout"]), stor_qt=dptr(keep["qt"]),
        c_shed=sys.c_shed, c_dump=sys.c_dump, c_viol=sys.c_viol, gamma_k=sys.gamma_k, kappa0=sys.kappa0,
        beta_v=sys.beta_v, beta_k=sys.beta_k, K=sys.K, Ptr=dptr(keep["Ptr"]), pi0=dptr(keep["pi0"]),
        phi=sys.phi, sigma=sys.sigma, sigma0=sys.sigma0, mu=dptr(keep["mu"]), xbar=dptr(keep["xbar"]),
        ncountry=int(getattr(sys, "ncountry", 0) or 0), country=iptr(keep["country"]),
    )
    return m, keep


# exported symbols and their signatures: (name, restype, argtypes)
SIGNATURES = [
    ("esm_version", C.c_int32, []),
    ("esm_last_error", C.c_char_p, []),
    ("esm_create", C.c_int32, [C.POINTER(EsmModel), C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    ("esm_destroy", C.c_int32, [C.c_void_p]),
    ("esm_upload_state", C.c_int32, [C.c_void_p, _dp, _dp, _dp]),
    ("esm_download_state", C.c_int32, [C.c_void_p, _dp, _dp, _dp]),
    ("esm_heuristic_state", C.c_int32, [C.c_void_p]),
    ("esm_energy", C.c_int32, [C.c_void_p, _dp, _dp]),
    ("esm_rebuild_residual", C.c_int32, [C.c_void_p, _dp]),
    ("esm_download_caches", C.c_int32, [C.c_void_p, _dp, _dp]),
    ("esm_scores", C.c_int32, [C.c_void_p, _dp, _dp]),
    ("esm_score_measure", C.c_int32, [C.c_void_p, C.c_int32]),
    ("esm_langevin", C.c_int32, [C.c_void_p, _dp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int32, _dp, _ip]),
    ("esm_langevin_noalloc", C.c_int32, [C.c_void_p, _dp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int32, _dp,
                                         _ip]),
    ("esm_sweeps", C.c_int32, [C.c_void_p, C.c_int32, C.POINTER(C.c_uint32), C.c_uint32, C.c_double,
                               C.POINTER(C.c_int64)]),
    ("esm_schedule", C.c_int32, [C.c_void_p, _ip, _ip]),
    ("esm_follow_capacities", C.c_int32, [C.c_void_p, _dp, C.c_double, _dp]),
    ("esm_weather_init", C.c_int32, [C.c_void_p, C.POINTER(C.c_uint32), C.c_uint32]),
    ("esm_weather_step", C.c_int32, [C.c_void_p, _dp, _dp, C.c_int32, C.c_int32, C.POINTER(C.c_uint32), C.c_uint32,
                                     _ip]),
    ("esm_weather_ffbs", C.c_int32, [C.c_void_p, _ip, C.POINTER(C.c_uint32), C.c_uint32]),
    ("esm_download_weather", C.c_int32, [C.c_void_p, _ip, _dp]),
    ("esm_replica_features", C.c_int32, [C.c_void_p, _dp, _dp]),
    ("esm_upload_weather", C.c_int32, [C.c_void_p, _ip, _dp]),
    ("esm_copy_replica", C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32]),
    ("esm_sweep_options", C.c_int32, [C.c_void_p, C.c_double, _dp, C.c_int32, C.c_double]),
    ("esm_storage_long_options", C.c_int32, [C.c_void_p, C.c_double, C.c_int32, C.c_double]),
    ("esm_storage_cross_options", C.c_int32, [C.c_void_p, C.c_double, C.c_int32]),
    ("esm_sweep_launch", C.c_int32, [C.c_void_p, C.c_int32]),
    ("esm_mlp_init", C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, _u32p]),
    ("esm_mlp_dims", C.c_int32, [C.c_void_p, _ip]),
    ("esm_mlp_params", C.c_int32, [C.c_void_p, _dp, _dp]),
    ("esm_mlp_optimizer", C.c_int32, [C.c_void_p, _dp, _dp]),
    ("esm_mlp_options", C.c_int32, [C.c_void_p, C.c_int32]),
    ("esm_mlp_forward", C.c_int32, [C.c_void_p, _dp, C.c_int64, _dp]),
    ("esm_mlp_gradient", C.c_int32, [C.c_void_p, _dp, _dp, _dp, C.c_int64, _dp, _dp]),
    ("esm_mlp_train", C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_double, _u32p, C.c_uint32, _dp]),
    ("esm_mlp_propose", C.c_int32, [C.c_void_p, _u32p, C.c_uint32, C.POINTER(C.c_int64), _dp, _dp, _dp]),
    ("esm_profile", C.c_int32, [C.c_void_p, C.c_int32]),
    ("esm_profile_read", C.c_int32, [C.c_void_p, C.c_int32, _dp, C.POINTER(C.c_int64)]),
    ("esm_batch_create", C.c_int32, [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_void_p)]),
    ("esm_batch_destroy", C.c_int32, [C.c_void_p]),
    ("esm_batch_sweeps", C.c_int32, [C.c_void_p, C.c_int32, _u32p, _u32p, C.c_double, C.POINTER(C.c_int64)]),
    ("esm_batch_counts", C.c_int32, [C.c_void_p, C.POINTER(C.c_int64)]),
    ("esm_sweep_block", C.c_int32, [C.c_void_p, C.c_int32]),
    ("esm_sweep_pad", C.c_int32, [C.c_void_p, C.c_int32]),
    ("esm_autotune", C.c_int32, [C.c_void_p, C.c_int32, _u32p, C.c_uint32, C.c_double, _ip, _dp, _ip]),
]
(KERN_ENERGY, KERN_SCORE, KERN_SWEEP, KERN_WEATHER, KERN_GEN, KERN_THETA, KERN_SOC, KERN_SPAIR, KERN_FOLLOW,
 KERN_UNIT, KERN_POT, KERN_CELL, KERN_XLOC, KERN_XLOCC, KERN_XPAIR, KERN_SMS, KERN_SLONG, KERN_MLPF, KERN_MLPP,
 KERN_MLPT, KERN_RANDOM, KERN_SXPAIR, KERN_SXLONG) = range(23)
NFAM, NSTAT, NWEIGHT = 15, 5, 9
BATCH_MAX = 64
MEASURES = {"utilization": 0, "physical": 1}   # ESM_MEASURE_*
# the plain random families and their default weights (plain_families, src/local.jl:1912-1915)
WEIGHT_NAMES = ("unit", "pair", "xpair", "xlocc", "xloc", "potential", "cellpair", "cell", "spair")
DEFAULT_WEIGHTS = (0.3, 0.25, 0.03, 0.08, 0.04, 0.1, 0.1, 0.1, 0.3)

_lib = None


def load() -> C.CDLL:
    """Load libesmhip.so (raises if it is missing: there is no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python esm-sampling_amd/build.py` "
                           "(or __graft_entry__.build()); the HIP library is the only implementation")
    lib = C.CDLL(str(LIB_PATH))
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class EsmError(RuntimeError):
    pass


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().esm_last_error()
        raise EsmError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
