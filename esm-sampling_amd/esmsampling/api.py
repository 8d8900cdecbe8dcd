"""Host-side mirror of the reference's per-sweep API over the C ABI of libesmhip.so.

The reference's boundary for this path is the Julia function API that ``run_replicas``
(src/replicas.jl:841) calls. This module keeps its names and argument meanings, for a batch of
replicas resident on one GPU (one ``Replicas`` object = one ``esm_handle``):

=================================  ============================================  ===================
reference (Julia)                  here                                          C entry point
=================================  ============================================  ===================
``energy_cost`` src/energy.jl:551  ``energy_cost(R)`` -> CostParts per replica   ``esm_energy``
``LocalCost`` src/local.jl:313     ``LocalCost(R)`` -> (r, F) caches             ``esm_rebuild_residual``
``capacity_scores`` replicas:329   ``capacity_scores(R)`` -> (sk, fk)            ``esm_scores``
``langevin_capacities`` :391       ``langevin_capacities(R, ...)``               ``esm_langevin``
=================================  ============================================  ===================

Errors raise ``EsmError`` with the library's message (the reference raises Julia exceptions).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib, synth
from ._lib import FP32, FP64, NK, NPARTS, NREN, NV, check, dptr, f64, iptr

PART_NAMES = ("capex", "op", "gen", "storage", "shed", "dump", "penalty", "co2")


class Replicas:
    """The device state of ``nrep`` replicas of one system (an ``esm_handle``).

    Arrays use the layout [field][node][hour][replica] (replica fastest).
    """

    def __init__(self, sys, nrep: int, precision: int = FP64):
        self.lib = _lib.load()
        self.sys = sys
        self.N, self.L, self.T, self.R = sys.N, sys.L, sys.T, int(nrep)
        self.precision = precision
        cm, keep = _lib.make_cmodel(sys)
        h = C.c_void_p()
        check(self.lib.esm_create(C.byref(cm), self.R, precision, C.byref(h)), "esm_create")
        self._h = h
        del keep

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            self.lib.esm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state -------------------------------------------------------------------------------
    def upload(self, v=None, k=None, x=None):
        N, T, R = self.N, self.T, self.R
        args = []
        keep = []
        for a, shape in ((v, (NV, N, T, R)), (k, (NK, N)), (x, (NREN, N, T, R))):
            if a is None:
                args.append(None)
                continue
            a = f64(a)
            if a.shape != shape:
                raise ValueError(f"expected shape {shape}, got {a.shape}")
            keep.append(a)
            args.append(dptr(a))
        check(self.lib.esm_upload_state(self._h, *args), "esm_upload_state")

    def heuristic(self):
        """The heuristic start state of every replica, set on the device from the current
        capacities and raw weather (``esm_heuristic_state``: zero angles, gas covering the net
        load, storage half full); the caches are rebuilt. The same values as
        ``synth.heuristic_state`` without staging the state on the host."""
        check(self.lib.esm_heuristic_state(self._h), "esm_heuristic_state")

    def download(self):
        N, T, R = self.N, self.T, self.R
        v = np.empty((NV, N, T, R))
        k = np.empty((NK, N))
        x = np.empty((NREN, N, T, R))
        check(self.lib.esm_download_state(self._h, dptr(v), dptr(k), dptr(x)), "esm_download_state")
        return v, k, x

    def caches(self):
        r = np.empty((self.N, self.T, self.R))
        F = np.empty((self.L, self.T, self.R))
        check(self.lib.esm_download_caches(self._h, dptr(r), dptr(F)), "esm_download_caches")
        return r, F


def energy_cost(R: Replicas, per_hour: bool = False):
    """CostParts of every replica ``[R, 8]`` (capex, op, gen, storage, shed, dump, penalty, co2);
    with ``per_hour`` also the operational cost of every step ``[T, R]``."""
    parts = np.empty((R.R, NPARTS))
    ph = np.empty((R.T, R.R)) if per_hour else None
    check(R.lib.esm_energy(R.handle, dptr(parts), dptr(ph) if per_hour else None), "esm_energy")
    return (parts, ph) if per_hour else parts


def LocalCost(R: Replicas):
    """Rebuild the caches from the node arrays and return ``(r [N,T,R], F [L,T,R], drift [R])``."""
    drift = np.empty(R.R)
    check(R.lib.esm_rebuild_residual(R.handle, dptr(drift)), "esm_rebuild_residual")
    r, F = R.caches()
    return r, F, drift


def rebuild_drift(R: Replicas):
    """Rebuild the caches from the node arrays and return the drift of the incremental residual
    cache per replica, max|r_cache − r| ``[R]`` (without downloading the caches)."""
    drift = np.empty(R.R)
    check(R.lib.esm_rebuild_residual(R.handle, dptr(drift)), "esm_rebuild_residual")
    return drift


def capacity_scores(R: Replicas):
    """``(sk [R, 6, N], fk [6, N])``: the sampled capacity score −β_v ∂C/∂k per replica and the capex force."""
    sk = np.empty((R.R, NK, R.N))
    fk = np.empty((NK, R.N))
    check(R.lib.esm_scores(R.handle, dptr(sk), dptr(fk)), "esm_scores")
    return sk, fk


def score_measure(R: Replicas, measure: str = "utilization"):
    """The reference measure of the capacity score (``measure``, src/run.jl:36): "utilization"
    (ŝ_k = −β_v ∂C/∂k) or "physical" (the dispatch of gas and storage units in MW: ŝ_k gains
    ``log_measure_grad`` = T/(k + κ0) inside the box, src/energy.jl:811-829)."""
    if measure not in _lib.MEASURES:
        raise ValueError(f'measure = "{measure}": "utilization" or "physical"')
    check(R.lib.esm_score_measure(R.handle, _lib.MEASURES[measure]), "esm_score_measure")


def langevin_capacities(R: Replicas, weights, xi, eta, kappa_s, kmin: float = 1.0, clip: float = 0.02,
                        box: bool = True):
    """One Langevin step of the shared capacities from the last scores: ``(knew [6, N], nclip)``."""
    w = f64(weights)
    xi = f64(xi)
    eta = f64(np.broadcast_to(eta, (NK,)))
    ks = f64(np.broadcast_to(kappa_s, (NK,)))
    if w.shape != (R.R,) or xi.shape != (NK, R.N):
        raise ValueError("weights must be [R] and xi [6, N]")
    knew = np.empty((NK, R.N))
    nclip = C.c_int32(0)
    check(R.lib.esm_langevin(R.handle, dptr(w), dptr(xi), dptr(eta), dptr(ks), float(kmin), float(clip), int(box),
                             dptr(knew), C.byref(nclip)), "esm_langevin")
    return knew, int(nclip.value)


def _key(key):
    k = (C.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    return k


FAMILY_NAMES = ("gen", "theta", "soc", "spair", "unit", "potential", "cell", "xloc", "xlocc", "xpair",
                "storage_ms", "slong", "slevel", "sxpair", "sxlong")


def sweep_options(R: Replicas, extra: float = 4.0, weights=None, atoms: bool = True, p_storage_ms: float = 0.5):
    """The `local_sweep!` options (src/local.jl:1927-1929): ``extra`` random moves per node and step,
    the weights of the plain families (``_lib.WEIGHT_NAMES`` order; None: the reference's),
    whether the atoms run, and the probability of the storage multiscale sweep that
    `operation_sweep!` runs after them (src/local.jl:2414-2418: 1/2)."""
    w = None if weights is None else f64(weights)
    if w is not None and w.shape != (_lib.NWEIGHT,):
        raise ValueError(f"weights must have {_lib.NWEIGHT} entries")
    check(R.lib.esm_sweep_options(R.handle, float(extra), dptr(w) if w is not None else None, int(bool(atoms)),
                                  float(p_storage_ms)), "esm_sweep_options")


def storage_long_options(R: Replicas, p: float = 1.0, nmoves: int = 32, hmin: float = 24.0):
    """The long-range storage sweep (`storage_long_sweep!`, src/storage_long.jl:245-298) that
    `operation_sweep!` runs with probability ``p`` (``storage_long``, src/run.jl:27-29): a level
    move and ``nmoves`` long-range plateaus (``storage_long_moves``) per storage unit with at least
    ``hmin`` hours (``storage_long_hours``)."""
    check(R.lib.esm_storage_long_options(R.handle, float(p), int(nmoves), float(hmin)), "esm_storage_long_options")


def storage_cross_options(R: Replicas, p_xpair: float = 0.1, nxlong: int = 8):
    """The cross-node storage pairs (src/storage_cross.jl): ``p_xpair`` the weight of the
    single-instant pair ``sxpair`` among the plain families of the local sweep (``storage_xpair``),
    ``nxlong`` the interval pairs ``sxlong`` per storage unit in the long-range sweep
    (``storage_xlong``); the units taking part have at least the long-range sweep's ``hmin``
    hours. The defaults here are the reference configuration's (src/run.jl:30-31); a new handle
    starts with both off, as `operation_sweep!`'s own keywords."""
    check(R.lib.esm_storage_cross_options(R.handle, float(p_xpair), int(nxlong)), "esm_storage_cross_options")


def sweep_launch(R: Replicas, merged=True):
    """The table-driven random families of a parity class in one launch (True or 1, the default)
    or one launch per family (False or 0: per-family kernel times; the same states and
    statistics), or merged with 2, 4, 8 or 16 waves per (step, 64 replicas) splitting every
    move's template entries (replica counts that are multiples of 64)."""
    mode = int(merged) if not isinstance(merged, bool) else (1 if merged else 0)
    check(R.lib.esm_sweep_launch(R.handle, mode), "esm_sweep_launch")


def operation_sweep(R: Replicas, nsweeps: int = 1, key=(1, 0), counter: int = 0, maxnats: float = 40.0):
    """``operation_sweep!`` / ``local_sweep!`` (src/local.jl:1927-2062, :2401-2425) of every replica,
    ``nsweeps`` times at fixed capacities and weather: the atoms, the random families, then (each
    with its probability) the storage multiscale and long-range storage sweeps. Returns the move
    statistics [15 families
    (``FAMILY_NAMES``), 5] (attempted, accepted,
    improper, point-cap / non-finite / Jacobian-rejected, ΔC evaluations)."""
    st = np.zeros(_lib.NFAM * _lib.NSTAT, dtype=np.int64)
    check(R.lib.esm_sweeps(R.handle, int(nsweeps), _key(key), int(counter) & 0xFFFFFFFF, float(maxnats),
                           st.ctypes.data_as(C.POINTER(C.c_int64))), "esm_sweeps")
    return st.reshape(_lib.NFAM, _lib.NSTAT)


def sweep_block(R: Replicas, block: int = 256):
    """Workgroup size of the one-wave random launches (64, 128 or 256); states and statistics do
    not depend on it."""
    check(R.lib.esm_sweep_block(R.handle, int(block)), "esm_sweep_block")


def sweep_pad(R: Replicas, pad: bool = True):
    """Every step of the one-wave random launches on whole waves (R rounded up to a multiple of 64,
    the extra lanes idle: one step per wave; the default) or R threads per step; states and
    statistics do not depend on it."""
    check(R.lib.esm_sweep_pad(R.handle, int(bool(pad))), "esm_sweep_pad")


AUTOTUNE_FORMS = (0, 1, 2, 4)
AUTOTUNE_BLOCKS = (64, 128, 256)


def autotune_candidates(nrep: int):
    """The (form, block, pad) of the autotuner's 12 candidates, in the library's order."""
    if nrep % 64 == 0:
        return [(AUTOTUNE_FORMS[c // 3], AUTOTUNE_BLOCKS[c % 3], 0) for c in range(12)]
    return [(AUTOTUNE_FORMS[(c % 6) // 3], AUTOTUNE_BLOCKS[c % 3], c // 6) for c in range(12)]


def _shape_key(R: Replicas) -> str:
    return f"N{R.N}_L{R.L}_T{R.T}_R{R.R}_fp{R.precision}"


def autotune(R: Replicas, nsweeps: int = 1, key=(7, 7), counter: int = 0, maxnats: float = 40.0,
             cache: str | None = None, retune: bool = False):
    """Pick the launch form of the random families (``sweep_launch``), their workgroup size
    (``sweep_block``) and padding (``sweep_pad``) for this shape by timing every candidate on the
    device (``esm_autotune``:
    each candidate runs ``nsweeps`` sweeps from the current state, which is restored afterwards,
    so the chain does not move). With ``cache`` (a JSON file), a shape tuned before takes the
    stored choice without timing (unless ``retune``), and a new choice is written to the file. Returns
    {"form", "block", "pad", "ms": {(form, block, pad): ms per sweep}, "cached"}."""
    import json
    import os
    sk = _shape_key(R)
    table = {}
    if cache and os.path.exists(cache):
        with open(cache) as f:
            table = json.load(f)
        if sk in table and not retune:
            form, block, pad = int(table[sk]["form"]), int(table[sk]["block"]), int(table[sk].get("pad", 0))
            sweep_launch(R, form)
            sweep_block(R, block)
            sweep_pad(R, pad)
            return {"form": form, "block": block, "pad": pad, "ms": {}, "cached": True}
    best = (C.c_int32 * 3)()
    ms = (C.c_double * 12)()
    nc = C.c_int32()
    check(R.lib.esm_autotune(R.handle, int(nsweeps), _key(key), int(counter) & 0xFFFFFFFF, float(maxnats), best, ms,
                             C.byref(nc)), "esm_autotune")
    cands = autotune_candidates(R.R)
    times = {cands[c]: float(ms[c]) for c in range(nc.value)}
    out = {"form": int(best[0]), "block": int(best[1]), "pad": int(best[2]), "ms": times, "cached": False}
    if cache:
        table[sk] = {"form": out["form"], "block": out["block"], "pad": out["pad"],
                     "ms": {f"{f}/{b}/{p}": t for (f, b, p), t in times.items()}}
        tmp = cache + ".tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, cache)
    return out


class Batch:
    """Replica sets of uneven size (network, horizon, replica count, precision) swept together as
    one HIP graph launch (``esm_batch_sweeps``): the same kernels, arguments and results as
    ``operation_sweep`` on each set, run as concurrent branches of one graph. The sets must not
    be used by other threads during a batched sweep; kernel profiling must be off on them. The
    batch refers to the sets' handles: close it before closing them."""

    def __init__(self, sets):
        self.sets = list(sets)
        if not 1 <= len(self.sets) <= _lib.BATCH_MAX:
            raise ValueError(f"a batch holds 1 to {_lib.BATCH_MAX} replica sets")
        self.lib = self.sets[0].lib
        hs = (C.c_void_p * len(self.sets))(*[r.handle.value for r in self.sets])
        b = C.c_void_p()
        check(self.lib.esm_batch_create(hs, len(self.sets), C.byref(b)), "esm_batch_create")
        self._b = b

    def sweeps(self, nsweeps: int = 1, keys=None, counters=None, maxnats: float = 40.0):
        """``operation_sweep`` of every set (set i under keys[i] and counters[i]); returns the
        statistics [n sets, 15 families, 5]."""
        n = len(self.sets)
        keys = [(1, 0)] * n if keys is None else list(keys)
        counters = [0] * n if counters is None else list(counters)
        if len(keys) != n or len(counters) != n:
            raise ValueError("one key and one counter per replica set")
        kk = (C.c_uint32 * (2 * n))(*[int(x) & 0xFFFFFFFF for k in keys for x in k])
        cc = (C.c_uint32 * n)(*[int(c) & 0xFFFFFFFF for c in counters])
        st = np.zeros(n * _lib.NFAM * _lib.NSTAT, dtype=np.int64)
        check(self.lib.esm_batch_sweeps(self._b, int(nsweeps), kk, cc, float(maxnats),
                                        st.ctypes.data_as(C.POINTER(C.c_int64))), "esm_batch_sweeps")
        return st.reshape(n, _lib.NFAM, _lib.NSTAT)

    def counts(self):
        """(graph launches, instantiations, in-place updates of the executable graph)."""
        c = (C.c_int64 * 3)()
        check(self.lib.esm_batch_counts(self._b, c), "esm_batch_counts")
        return tuple(int(x) for x in c)

    def close(self):
        if self._b:
            self.lib.esm_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def follow_capacities(R: Replicas, knew, kmin: float = 1.0):
    """``follow_capacities!`` (src/replicas.jl:448-472): move the shared capacities to ``knew`` [6, N]
    with the dispatch following on every replica; returns the committed capacities [6, N]."""
    kn = f64(knew)
    if kn.shape != (NK, R.N):
        raise ValueError("knew must be [6, N]")
    kout = np.empty((NK, R.N))
    check(R.lib.esm_follow_capacities(R.handle, dptr(kn), float(kmin), dptr(kout)), "esm_follow_capacities")
    return kout


def schedule(R: Replicas):
    """(colour [N], number of colours) of the angle rounds."""
    col = np.zeros(R.N, dtype=np.int32)
    nc = C.c_int32()
    check(R.lib.esm_schedule(R.handle, iptr(col), C.byref(nc)), "esm_schedule")
    return col, int(nc.value)


def sample_weather(R: Replicas, key=(1, 0), counter: int = 0):
    """Exact prior draw of every replica's weather (``sample_weather``, src/weather.jl:461-481):
    regimes, VAR anomalies and the raw weather x; the caches are rebuilt."""
    check(R.lib.esm_weather_init(R.handle, _key(key), int(counter)), "esm_weather_init")


def weather_process_step(R: Replicas, theta, win_prob, window: int = 60, compensate: bool = True, key=(1, 0),
                         counter: int = 0):
    """One weather step of every replica (``replica_weather!`` / ``weather_process_step`` and
    ``compensate_weather!``, src/replicas.jl:179-236): pCN of angle theta[j], a regime window of
    length ``window`` with probability win_prob[j]. Returns the move flags [R] (0 none, 1 pCN, 2 window)."""
    th = f64(np.broadcast_to(theta, (R.R,)))
    wp = f64(np.broadcast_to(win_prob, (R.R,)))
    moved = np.zeros(R.R, dtype=np.int32)
    check(R.lib.esm_weather_step(R.handle, dptr(th), dptr(wp), int(window), int(compensate), _key(key), int(counter),
                                 iptr(moved)), "esm_weather_step")
    return moved


def weather_ffbs(R: Replicas, mask=None, key=(1, 0), counter: int = 0):
    """The centred FFBS of the regimes (``replica_ffbs``: ``fine_regime_step!`` by ``ffbs``,
    src/weather.jl:160-174) of the replicas with ``mask[j]`` set (None: all): every regime path is
    redrawn from p(m | z); the anomalies, the weather and the dispatch are unchanged."""
    mk = None if mask is None else np.ascontiguousarray(np.broadcast_to(mask, (R.R,)), dtype=np.int32)
    check(R.lib.esm_weather_ffbs(R.handle, iptr(mk) if mk is not None else None, _key(key), int(counter)),
          "esm_weather_ffbs")


def weather_state(R: Replicas):
    """(m [T, R] regimes, z [3N, T, R] anomalies) of every replica."""
    m = np.empty((R.T, R.R), dtype=np.int32)
    z = np.empty((NREN * R.N, R.T, R.R))
    check(R.lib.esm_download_weather(R.handle, iptr(m), dptr(z)), "esm_download_weather")
    return m, z


def replica_feature_names(sys) -> list:
    """Names of the replica weather features (``replica_feature_names``, src/replicas.jl:274-276):
    the annual mean load, the capacity factors of the weather carriers, the residual load's mean,
    peak and worst 7-step mean (GW), the regime frequencies and the number of regime switches."""
    K = int(getattr(sys, "K", 1) or 1)
    return (["load_GW"] + [f"cf_{c}" for c in synth.CARRIERS[:NREN]] +
            ["resload_mean_GW", "resload_peak_GW", "resload_7step_GW"] + [f"regime_{k}" for k in range(1, K + 1)] +
            ["regime_switches"])


def replica_features(R: Replicas, kref) -> np.ndarray:
    """The weather features [R, nf] of every replica (``replica_features``, src/replicas.jl:270-273)
    at the reference capacities ``kref`` [6, N] (the renewable fields are read), in the order of
    ``replica_feature_names``."""
    kref = f64(kref)
    if kref.shape != (len(synth.CARRIERS), R.N):
        raise ValueError(f"kref must be [{len(synth.CARRIERS)}, N]")
    nf = len(replica_feature_names(R.sys))
    out = np.empty((R.R, nf))
    check(R.lib.esm_replica_features(R.handle, dptr(kref), dptr(out)), "esm_replica_features")
    return out


def set_weather(R: Replicas, m, z):
    m = np.ascontiguousarray(m, dtype=np.int32)
    z = f64(z)
    if m.shape != (R.T, R.R) or z.shape != (NREN * R.N, R.T, R.R):
        raise ValueError("m must be [T, R] and z [3N, T, R]")
    check(R.lib.esm_upload_weather(R.handle, iptr(m), dptr(z)), "esm_upload_weather")


def copy_replica(src: Replicas, js: int, dst: Replicas, jd: int):
    """Copy replica ``js`` of ``src`` into replica ``jd`` of ``dst`` on the device (operation, weather
    and capacities; ``dst``'s caches are rebuilt): ``copy(s)`` of the lag probe, src/replicas.jl:736."""
    check(src.lib.esm_copy_replica(src.handle, int(js), dst.handle, int(jd)), "esm_copy_replica")


def weather_schedule(it: int, j: int, M: int, theta: float, burst_every: int, phi: float):
    """The weather angle of replica j (0-based) at step ``it`` and whether it is a burst
    (src/replicas.jl:571-576): bursts every ``burst_every`` steps, staggered by
    (j)·⌊burst_every/M⌋; a burst composes the continuous angle θ with φ."""
    import math
    nb = burst_every
    burst = nb > 0 and (it + j * (nb // M)) % nb == 0
    if burst:
        th = math.acos(max(-1.0, min(1.0, math.cos(theta) * math.cos(phi)))) if theta > 0 else phi
    else:
        th = theta
    return th, burst


def uniform_weights(active):
    """Weights 1/M_active over the active replicas (all if none is active), src/replicas.jl:310-314."""
    a = np.asarray(active, dtype=bool)
    if not a.any():
        a = np.ones_like(a)
    w = a.astype(np.float64)
    return w / w.sum()


def tilt_weights(lam: float, phis, active):
    """Self-normalised weights exp(λΦ_j) over the active replicas, src/replicas.jl:303-308."""
    a = np.asarray(active, dtype=bool)
    if not a.any():
        a = np.ones_like(a)
    x = np.where(a, lam * np.asarray(phis, dtype=np.float64), -np.inf)
    w = np.exp(x - x.max())
    return w / w.sum()


class LearnedProposal:
    """The learned dispatch-block proposal of a replica set (``esm_mlp_*``): an MLP proposes the
    gas utilization of one node over a block of ``block_hours`` hours from the block's context
    (the requirement the unit meets, the node's capacity factors and capacities); every proposal
    is accepted or rejected by the exact Metropolis–Hastings ratio (FP64), so the target is
    unchanged and the plain line moves remain both fallback and reference. Trained on the sampler's own states by
    maximum likelihood (Adam). No counterpart in the reference."""

    def __init__(self, R: Replicas, block_hours: int = 24, ncomp: int = 4, key=(7, 7)):
        self.R = R
        check(R.lib.esm_mlp_init(R.handle, int(block_hours), int(ncomp), _key(key)), "esm_mlp_init")
        d = np.zeros(7, dtype=np.int32)
        check(R.lib.esm_mlp_dims(R.handle, d.ctypes.data_as(C.POINTER(C.c_int32))), "esm_mlp_dims")
        self.B, self.K, self.din, self.dh, self.dout, self.nb, self.nparam = (int(q) for q in d)
        self.n_inst = R.sys.N * self.nb * R.R

    def params(self, set_to=None):
        """The FP32 master parameters (per layer: Wt [out][in], then the bias); with ``set_to``, set them first."""
        out = np.zeros(self.nparam)
        src = None if set_to is None else f64(set_to)
        if src is not None and src.shape != (self.nparam,):
            raise ValueError(f"{self.nparam} parameters expected")
        check(self.R.lib.esm_mlp_params(self.R.handle, dptr(out), dptr(src) if src is not None else None), "esm_mlp_params")
        return out

    def options(self, fused: bool = True):
        """The forward pass fused into one kernel per 128-row tile (default) or one kernel per layer."""
        check(self.R.lib.esm_mlp_options(self.R.handle, 1 if fused else 0), "esm_mlp_options")

    def optimizer(self, set_to=None):
        """Adam's state [m (nparam)][v (nparam)][t] (``esm_mlp_optimizer``); with ``set_to``, set it first."""
        out = np.zeros(2 * self.nparam + 1)
        src = None if set_to is None else f64(set_to)
        if src is not None and src.shape != (2 * self.nparam + 1,):
            raise ValueError(f"{2 * self.nparam + 1} values expected")
        check(self.R.lib.esm_mlp_optimizer(self.R.handle, dptr(out), dptr(src) if src is not None else None),
              "esm_mlp_optimizer")
        return out

    def layers(self, p=None):
        """[(Wt [out][in], b [out])] of the four layers, from ``p`` (or the current parameters)."""
        p = self.params() if p is None else p
        dims = [(self.din, self.dh), (self.dh, self.dh), (self.dh, self.dh), (self.dh, self.dout)]
        out, o = [], 0
        for fi, fo in dims:
            W = p[o:o + fi * fo].reshape(fo, fi)
            o += fi * fo
            b = p[o:o + fo]
            o += fo
            out.append((W, b))
        return out

    def forward(self, x):
        x = f64(x)
        out = np.zeros((x.shape[0], self.dout))
        check(self.R.lib.esm_mlp_forward(self.R.handle, dptr(x), x.shape[0], dptr(out)), "esm_mlp_forward")
        return out

    def gradient(self, x, y, mu0):
        x, y, mu0 = f64(x), f64(y), f64(mu0)
        g = np.zeros(self.nparam)
        loss = C.c_double()
        check(self.R.lib.esm_mlp_gradient(self.R.handle, dptr(x), dptr(y), dptr(mu0), x.shape[0], dptr(g),
                                          C.byref(loss)), "esm_mlp_gradient")
        return float(loss.value), g

    def train(self, steps: int, batch: int = 16384, lr: float = 1e-3, key=(9, 9), counter: int = 0):
        """``steps`` Adam steps on minibatches of the current states; the mean loss of every step."""
        loss = np.zeros(max(int(steps), 1))
        check(self.R.lib.esm_mlp_train(self.R.handle, int(steps), int(batch), float(lr), _key(key),
                                       int(counter) & 0xFFFFFFFF, dptr(loss)), "esm_mlp_train")
        return loss[:int(steps)]

    def propose(self, key=(11, 11), counter: int = 0, trace: bool = False):
        """One learned move of every (node, block, replica); returns the statistics [5] (attempted,
        accepted, 0, 0, network evaluations) and, with ``trace``, the network outputs, proposals and
        MH log-ratios per instance ((node·nb + block)·R + replica)."""
        st = np.zeros(5, dtype=np.int64)
        out = yp = lr = None
        if trace:
            out = np.zeros((self.n_inst, self.dout))
            yp = np.zeros((self.n_inst, self.B))
            lr = np.zeros(self.n_inst)
        check(self.R.lib.esm_mlp_propose(self.R.handle, _key(key), int(counter) & 0xFFFFFFFF,
                                         st.ctypes.data_as(C.POINTER(C.c_int64)), dptr(out) if trace else None,
                                         dptr(yp) if trace else None, dptr(lr) if trace else None), "esm_mlp_propose")
        return (st, out, yp, lr) if trace else st
