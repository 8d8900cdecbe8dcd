"""Records, outputs and snapshots of the replica driver.

The reference writes them with HDF5 (``write_replica_outputs``, ``write_replica_summary``,
``write_replica_snapshot`` / ``read_replica_snapshot``, src/replicas.jl:961-1167; layouts in
docs/system/data.md §3 and §5). HDF5 is not available to this package, so the same datasets
are written as NumPy ``.npz`` archives whose keys are the reference's dataset paths
(``replicas/1/ktot``, ``chains/1/eta_w``, …) and whose matrices keep its layout, one column
per step (quantity × step). Files are replaced atomically (written to a temporary name in the
same directory, then renamed), as the reference's ``atomic_write`` does.

Snapshot layout (version 1, ``driver = "replicas"``): ``phase``, ``it`` (replica steps completed),
``where``, ``restarts``, ``written``, ``config/seed``, ``nchains``; per chain c (1-based)
``chains/<c>/start``, ``replicas``, ``Q`` (monitored trace, quantity × step), ``eta_w``, ``since``,
``window``, ``active``, ``trace/<key>``, ``probe/<key>``. The replica states are stored per chain
in the device layout ``[field][node][hour][replica]`` (``state/v``, ``state/x``, ``state/k``, the
weather ``state/m`` and ``state/z``) instead of one group per replica; the driver's own state
for an exact continuation is stored beside them (``driver/gbuf``, ``driver/hist``, the step
adaptation's pending gaps ``driver/gaps``, the Langevin normals' generator state ``driver/rng``,
the probe counter, the learned proposal's parameters and Adam state ``learned/adam``), so that a
run resumed anywhere, the warmup included, continues as the uninterrupted run.
"""
from __future__ import annotations

import json
import os
import tempfile
import time

import numpy as np

from . import synth

SNAPSHOT_VERSION = 1
EXIT_CHECKPOINTED = 75   # src/checkpoint.jl:113: a run stopped by the stop file after a snapshot


def _atomic_savez(path: str, arrays: dict):
    d = os.path.dirname(os.path.abspath(path)) or "."
    os.makedirs(d, exist_ok=True)
    fd, tmp = tempfile.mkstemp(prefix=".tmp_", suffix=".npz", dir=d)
    os.close(fd)
    try:
        with open(tmp, "wb") as f:
            np.savez(f, **arrays)
            f.flush()
            os.fsync(f.fileno())
        os.replace(tmp, path)
    except BaseException:
        if os.path.exists(tmp):
            os.remove(tmp)
        raise


def record_matrix(v):
    """Matrix of a record list, one column per step (``record_matrix``, src/replicas.jl:958)."""
    if len(v) == 0:
        return np.zeros((0, 0))
    return np.stack([np.atleast_1d(np.asarray(x, dtype=np.float64)) for x in v], axis=1)


# the replica records of one chain per step (keys of replicas/<c>/..., docs/system/data.md §3)
REPLICA_KEYS = ("ktot", "capex", "op", "shed", "penalty", "pfcost", "weights", "active", "since", "phi", "moved",
                "rs", "fk", "eta_w", "nclip", "time", "learned", "features")
PROBE_KEYS = ("it", "gap", "rs_eq", "rs_in", "fk", "cost", "eta_before", "eta_after")


def replica_step_record(k, parts, wts, active, since, moved, rs, fk, eta_w, nclip, wall, loop, learned):
    """One step's replica record (every array a copy, as the driver updates its own in place):
    carrier totals of the shared capacities, capex, per replica
    opex, shed, penalty and penalty-free cost (capex + generation + storage), the drift weights,
    active flags, steps since the last burst, Φ (NaN: no declared tilt functional), whether the
    weather moved, the interior means of r s̄ and β_k ∂K, the Langevin steps, the clipped drifts,
    (wall time, time of the replica loop) and the learned proposal's (attempted, accepted)."""
    return {"ktot": np.asarray(k).sum(axis=1), "capex": [float(parts[0, 0])], "op": parts[:, 1].copy(),
            "shed": parts[:, 4].copy(), "penalty": parts[:, 6].copy(),
            "pfcost": parts[:, 0] + parts[:, 2] + parts[:, 3], "weights": np.array(wts, dtype=np.float64),
            "active": np.array(active, dtype=bool), "since": np.minimum(np.asarray(since), 2 ** 31).astype(float),
            "phi": np.full(len(wts), np.nan), "moved": np.array(moved, dtype=np.float64), "rs": np.array(rs),
            "fk": np.array(fk), "eta_w": np.array(eta_w, dtype=np.float64), "nclip": [float(nclip)],
            "time": [float(wall), float(loop)], "learned": np.array(learned, dtype=np.float64)}


def probe_record(it, j, adapted, pr, eta_before, eta_after):
    """One probe's record (replicas/<c>/probe/...)."""
    return {"it": [float(it), float(j + 1), float(bool(adapted))], "gap": pr["gap"], "rs_eq": pr["rs_eq"],
            "rs_in": pr["rs_in"], "fk": pr["fk"],
            "cost": [pr["op_in"], pr["op_eq"], pr["shed_in"], pr["shed_eq"]],
            "eta_before": np.array(eta_before), "eta_after": np.array(eta_after)}   # copies: the driver
                                                                                  # adapts its η in place


def _columns(recs, keys):
    return {k: record_matrix([r[k] for r in recs]) for k in keys if recs and k in recs[0]}


def write_replica_outputs(path, cfg, sys, ctl, rrecs, precs, restarts=0, resumed_at=-1):
    """``write_replica_outputs`` (src/replicas.jl:961-986): the replica and probe records of every
    chain (``replicas/<c>/features``: the replicas' weather features stacked replica by replica),
    the carriers, the feature names, M, κ_s, the tilt, the weather transition matrix, β and the
    restarts."""
    from . import api
    out = {"replicas/carriers": np.array(synth.CARRIERS),
           "replicas/feature_names": np.array(api.replica_feature_names(sys), dtype=str),
           "replicas/M": np.array(ctl.M), "replicas/kappa_s": np.asarray(ctl.kappa_s),
           "replicas/tilt": np.array([0.0, np.nan, np.nan]), "run/restarts": np.array(restarts),
           "run/resumed_at_sweep": np.array(resumed_at), "calibration/beta": np.array([cfg.beta_k, cfg.beta_v])}
    if getattr(sys, "Ptr", None) is not None:
        out["weather/transition"] = np.asarray(sys.Ptr, dtype=np.float64)
    for c, recs in enumerate(rrecs):
        for k, A in _columns(recs, REPLICA_KEYS).items():
            out[f"replicas/{c + 1}/{k}"] = A
        for k, A in _columns(precs[c], PROBE_KEYS).items():
            out[f"replicas/{c + 1}/probe/{k}"] = A
    _atomic_savez(path, out)
    return path


def summary_path(out: str) -> str:
    """``summary_replicas.txt`` next to ``out`` (``replica_summary_path``, src/replicas.jl:989)."""
    return os.path.join(os.path.dirname(os.path.abspath(out)), "summary_replicas.txt")


def write_replica_summary(path, ctl, rrecs, precs, kinds, nwarm: int, passed: bool, wall: float, feature_names=()):
    """``write_replica_summary`` (src/replicas.jl:991-1047) over the production steps (all steps
    if there are none) per chain: carrier totals (mean, sd, last), the last Langevin steps, the
    drift and the probes' equilibrated scores and gaps in units of β_k ∂K, capex, penalty-free
    cost, opex, shed, the mean number of active replicas, and the replica weather features
    (mean and sd over all replicas and steps)."""
    fmt = lambda x: f"{x:.6g}"
    cars = synth.CARRIERS

    def row(name, v):
        return name.ljust(26) + " ".join(fmt(x).rjust(13) for x in v)

    def fmean(a):
        a = np.asarray(a, dtype=np.float64)
        a = a[np.isfinite(a)]
        return a.mean() if a.size else np.nan

    lines = [f"# replica–Langevin run: chains {len(rrecs)}, replicas {ctl.M}, steps {len(rrecs[0]) if rrecs else 0} "
             f"(warmup {nwarm}), λ 0, wall {round(wall)} s, convergence {'passed' if passed else 'not passed'}"]
    for c, recs in enumerate(rrecs):
        n = len(recs)
        w = slice(nwarm, n) if nwarm < n else slice(0, n)
        cols = {k: v[:, w] for k, v in _columns(recs, REPLICA_KEYS).items()}
        if not recs:
            continue
        lines.append("")
        lines.append(f"## chain {c + 1} ({kinds[c]}), steps {w.start + 1}–{n}")
        lines.append("".ljust(26) + " ".join(x.rjust(13) for x in cars))
        K = cols["ktot"] / 1e3
        lines.append(row("k total GW, mean", K.mean(axis=1)))
        lines.append(row("k total GW, sd", K.std(axis=1, ddof=1) if K.shape[1] > 1 else np.zeros(len(cars))))
        lines.append(row("k total GW, last", K[:, -1]))
        lines.append(row("eta_w, last", recs[-1]["eta_w"]))
        rs, fk = cols["rs"], cols["fk"]
        lines.append(row("drift r s̄/β_k∂K, mean", [fmean(rs[f] / fk[f]) for f in range(len(cars))]))
        pr = precs[c]
        if pr:
            P = _columns(pr, PROBE_KEYS)
            sel = np.flatnonzero(P["it"][0] > nwarm) if nwarm < n else np.arange(P["it"].shape[1])
            if sel.size == 0:
                sel = np.arange(P["it"].shape[1])
            eq, fkp, gp = P["rs_eq"][:, sel], P["fk"][:, sel], P["gap"][:, sel]
            lines.append(row("probe r ŝ_eq/β_k∂K", [fmean(eq[f] / fkp[f]) for f in range(eq.shape[0])]))
            lines.append(row("probe gap/β_k∂K", [fmean(gp[f]) for f in range(gp.shape[0])]))
            lines.append("probes".ljust(26) + str(sel.size))
        for key, lab in (("capex", "capex €/a"), ("pfcost", "pf cost €/a (replicas)"), ("op", "opex €/a (replicas)"),
                         ("shed", "shed MWh (replicas)")):
            A = cols[key]
            lines.append(lab.ljust(26) + "mean " + fmt(A.mean()) + "  sd " + fmt(A.std(ddof=1) if A.size > 1 else 0.0))
        lines.append("active replicas".ljust(26) + "mean " + fmt(cols["active"].sum(axis=0).mean()) + f" of {ctl.M}")
        nf = len(feature_names)
        if nf and "features" in cols and cols["features"].size:
            X = cols["features"].T.reshape(-1, nf)
            lines.append("replica weather features (all replicas and steps): mean, sd")
            for i, name in enumerate(feature_names):
                sd = X[:, i].std(ddof=1) if X.shape[0] > 1 else 0.0
                lines.append("  " + name.ljust(24) + fmt(X[:, i].mean()).rjust(13) + fmt(sd).rjust(13))
    d = os.path.dirname(os.path.abspath(path)) or "."
    fd, tmp = tempfile.mkstemp(prefix=".tmp_", suffix=".txt", dir=d)
    with os.fdopen(fd, "w") as f:
        f.write("\n".join(lines) + "\n")
    os.replace(tmp, path)
    return path


def _ragged(lists):
    """Per replica lists of equal-shape arrays as (values [M, n, size], lengths [M], item shape)."""
    lens = np.array([len(x) for x in lists], dtype=np.int64)
    n = int(lens.max()) if lens.size else 0
    items = [np.asarray(v, dtype=np.float64) for x in lists for v in x]
    shape = items[0].shape if items else (1,)
    size = int(np.prod(shape)) if shape else 1
    out = np.full((len(lists), max(n, 1), size), np.nan)
    for j, x in enumerate(lists):
        for i, v in enumerate(x):
            out[j, i] = np.asarray(v, dtype=np.float64).ravel()
    return out, lens, np.array(shape, dtype=np.int64)


def _unragged(vals, lens, shape):
    shape = tuple(int(q) for q in shape)
    return [[vals[j, i].reshape(shape).copy() for i in range(int(lens[j]))] for j in range(len(lens))]


def write_snapshot(path, cfg, it, restarts, kinds, sets, traces, rrecs, precs, rngs, nprobe, learned=None,
                   prod_s=0.0):
    """Snapshot of the replica driver after `it` replica steps (``write_replica_snapshot``,
    src/replicas.jl:1082-1124, in the layout of this module's docstring)."""
    where = f"replicas: step {it} of {cfg.mlda_warmup + cfg.nsamples}" + (
        " (production)" if it > cfg.mlda_warmup else " (warmup)")
    out = {"version": np.array(SNAPSHOT_VERSION), "driver": np.array("replicas"), "phase": np.array("replicas"),
           "it": np.array(it), "where": np.array(where), "restarts": np.array(restarts), "written": np.array(time.time()),
           "config/seed": np.array(cfg.seed), "nchains": np.array(len(sets)), "run/production_s": np.array(prod_s)}
    for c, R in enumerate(sets):
        g = f"chains/{c + 1}"
        v, k, x = R.dev.download()
        m, z = api_weather(R.dev)
        out.update({f"{g}/start": np.array(kinds[c]), f"{g}/replicas": np.array(R.dev.R), f"{g}/eta_w": R.eta_w,
                    f"{g}/since": R.since, f"{g}/window": R.window.astype(np.int64),
                    f"{g}/active": R.active.astype(np.int64), f"{g}/Q": record_matrix(traces[c]),
                    f"{g}/state/v": v, f"{g}/state/x": x, f"{g}/state/k": k, f"{g}/state/kraw": R.k,
                    f"{g}/state/m": m, f"{g}/state/z": z,
                    f"{g}/driver/rng": np.array(rng_state_json(rngs[c].bit_generator.state)),
                    f"{g}/driver/nprobe": np.array(nprobe[c])})
        # the step adaptation's gaps collected since its last update (probe_avg > 1)
        out[f"{g}/driver/gaps"] = (np.array(R.gaps, dtype=np.float64).reshape(len(R.gaps), -1) if R.gaps
                                   else np.zeros((0, 0)))
        gv, gl, gs = _ragged(R.gbuf)
        hv, hl, hs = _ragged(R.hist)
        out.update({f"{g}/driver/gbuf": gv, f"{g}/driver/gbuf_len": gl, f"{g}/driver/gbuf_shape": gs,
                    f"{g}/driver/hist": hv, f"{g}/driver/hist_len": hl, f"{g}/driver/hist_shape": hs})
        for key, A in _columns(rrecs[c], REPLICA_KEYS + ("it",)).items():
            out[f"{g}/trace/{key}"] = A
        if rrecs[c]:
            out[f"{g}/trace/k"] = record_matrix([np.asarray(r["k"]).ravel() for r in rrecs[c]])
        for key, A in _columns(precs[c], PROBE_KEYS).items():
            out[f"{g}/probe/{key}"] = A
        if learned is not None and learned[c] is not None:
            L = learned[c]
            out.update({f"{g}/learned/params": L.net.params(), f"{g}/learned/adam": L.net.optimizer(),
                        f"{g}/learned/ntrain": np.array(L.ntrain),
                        f"{g}/learned/counts": np.array([L.att, L.acc]), f"{g}/learned/loss": np.asarray(L.loss),
                        f"{g}/learned/wall": np.array(L.wall)})
    _atomic_savez(path, out)
    return path


def rng_state_json(state) -> str:
    """A numpy bit generator's state as JSON (its arrays as tagged lists with their dtype)."""
    def enc(o):
        if isinstance(o, np.ndarray):
            return {"__ndarray__": o.tolist(), "dtype": str(o.dtype)}
        if isinstance(o, dict):
            return {k: enc(v) for k, v in o.items()}
        if isinstance(o, np.integer):
            return int(o)
        return o
    return json.dumps(enc(state))


def rng_state_from_json(text: str):
    def dec(o):
        if isinstance(o, dict):
            if "__ndarray__" in o:
                return np.array(o["__ndarray__"], dtype=o["dtype"])
            return {k: dec(v) for k, v in o.items()}
        return o
    return dec(json.loads(text))


def api_weather(dev):
    from . import api
    return api.weather_state(dev)


def read_snapshot(path, cfg):
    """The snapshot at `path` as a dict (None if there is none); checks the driver, the version,
    the seed and the chain count (``read_replica_snapshot``, src/replicas.jl:1133-1167)."""
    if not path or not os.path.isfile(path):
        return None
    with np.load(path, allow_pickle=False) as f:
        d = {k: f[k] for k in f.files}
    if "driver" not in d or str(d["driver"]) != "replicas":
        raise ValueError(f"snapshot {path} is not a snapshot of the replica driver")
    if int(d["version"]) != SNAPSHOT_VERSION:
        raise ValueError(f"snapshot {path}: format version {int(d['version'])}, the code reads {SNAPSHOT_VERSION}")
    if int(d["config/seed"]) != cfg.seed:
        raise ValueError(f"snapshot {path}: seed {int(d['config/seed'])}, the configuration has {cfg.seed}")
    if int(d["nchains"]) != cfg.nchains:
        raise ValueError(f"snapshot {path}: {int(d['nchains'])} chains, the configuration has {cfg.nchains}")
    return d


def chain_records(d, c):
    """(trace, replica records, probe records) of chain c (0-based) of a snapshot dict."""
    g = f"chains/{c + 1}"

    def cols(prefix, keys):
        mats = {k: d[f"{prefix}/{k}"] for k in keys if f"{prefix}/{k}" in d}
        if not mats:
            return []
        n = next(iter(mats.values())).shape[1]
        return [{k: A[:, i] for k, A in mats.items()} for i in range(n)]

    Q = d.get(f"{g}/Q", np.zeros((0, 0)))
    trace = [Q[:, i] for i in range(Q.shape[1])]
    recs = cols(f"{g}/trace", REPLICA_KEYS + ("it",))
    kf = d.get(f"{g}/trace/k")
    nk = d[f"{g}/state/k"].shape
    for i, r in enumerate(recs):
        r["it"] = int(r["it"][0])
        r["active"] = r["active"].astype(bool)
        r["nclip"] = int(r["nclip"][0])
        if kf is not None:
            r["k"] = kf[:, i].reshape(nk)
    return trace, recs, cols(f"{g}/probe", PROBE_KEYS)


def restore_driver_lists(d, c):
    g = f"chains/{c + 1}"
    gbuf = _unragged(d[f"{g}/driver/gbuf"], d[f"{g}/driver/gbuf_len"], d[f"{g}/driver/gbuf_shape"])
    hist = _unragged(d[f"{g}/driver/hist"], d[f"{g}/driver/hist_len"], d[f"{g}/driver/hist_shape"])
    return gbuf, hist
