#!/usr/bin/env python3
"""Throughput of the replica sampler's per-sweep path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--replicas R] [--hours T] [--fp64]

A step is one replica step of the plain GPU sampler over all replicas (``replica_step!``,
src/replicas.jl:676-705, with ``replica_inner = 1``): the pieces that the HIP library provides,
each over every (replica, hour) of the synthetic 40-node / 70-line system. Inputs are synthetic
(seeded) and resident in HBM before the timed region. Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "esm-sampling_amd"))

# MI355X published peaks
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--replicas", type=int, default=64)
    ap.add_argument("--hours", type=int, default=8760)
    ap.add_argument("--fp64", action="store_true", help="FP64 device state (default FP32 terms, FP64 totals)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    return ap.parse_args()


def energy_bytes_per_unit(N: int, L: int, es: int) -> int:
    """Algorithmic HBM bytes of the energy kernel per (hour, replica): reads θ, gas, 2 SoC, 3
    capacity factors per node (7N), writes r per node and F per line, plus 6 FP64 step parts."""
    return es * (7 * N + N + L) + 6 * 8


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: every box of this build has a single MI355X "
              f"(asked for --gpus {args.gpus}); the multi-GPU design is described, not run", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from esmsampling import FP32, FP64, _lib, api, synth

    T, R = args.hours, args.replicas
    prec = FP64 if args.fp64 else FP32
    s = synth.make_system(N=40, L=70, T=T, dt=8760.0 / T, seed=1)
    m, z = synth.sample_weather_host(s, R, seed=2)
    x = synth.weather_from_z(s, z)
    v = synth.heuristic_state(s, x)
    k = s.k0.copy()
    dev = api.Replicas(s, R, prec)
    dev.upload(v=v, k=k, x=x)
    w = api.uniform_weights(np.ones(R, dtype=bool))
    rng = np.random.default_rng(5)
    eta = np.full(6, 5e-4)
    ks = np.array([1000.0, 500.0, 200.0, 300.0, 100.0, 100.0])

    def step():
        api.energy_cost(dev)
        api.capacity_scores(dev)
        xi = rng.standard_normal((6, s.N))
        api.langevin_capacities(dev, w, xi, eta, ks, kmin=1.0, clip=0.02, box=True)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    dev.lib.esm_profile(dev.handle, 1)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    import ctypes as C
    tot = C.c_double()
    cnt = C.c_int64()
    dev.lib.esm_profile_read(dev.handle, _lib.KERN_ENERGY, C.byref(tot), C.byref(cnt))
    dev.lib.esm_profile(dev.handle, 0)
    ms_step = 1e3 * (t1 - t0) / args.steps
    value = args.steps / (t1 - t0)
    es = 8 if prec == FP64 else 4
    unit_bytes = energy_bytes_per_unit(s.N, s.L, es)
    kern_ms = tot.value / max(cnt.value, 1)
    achieved = unit_bytes * T * R / (kern_ms * 1e-3) / 1e9
    out = {
        "metric": "replica steps/s (sweeps/s of the replica sampler)",
        "value": value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == FP64 else "f32 terms, f64 totals",
        "data": "synthetic (seeded 40-node/70-line system, VAR weather prior draws)",
        "config": {"workload": f"replica sampler, 40 nodes / 70 lines, {T} steps, {R} replicas; "
                               "step = energy + flows + residual caches, capacity score, Langevin",
                   "replicas": R, "hours": T, "nodes": s.N, "lines": s.L},
        "roofline": {"kernel": "k_energy", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel_ms": kern_ms, "bytes_per_unit": unit_bytes, "units_per_launch": T * R},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(s, v, k, x, R, T, args.cpu_seconds)
    dev.close()
    print(json.dumps(out), flush=True)


def cpu_baseline(s, v, k, x, R, T, budget_s):
    """The FP64 C oracle on the host cores: energy + scores on a bounded sample of replicas,
    scaled to whole steps over all replicas."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np
    import oracle
    o = oracle.Oracle(s)
    nr = 1
    vs, xs = np.ascontiguousarray(v[..., :nr]), np.ascontiguousarray(x[..., :nr])
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < budget_s:
        o.energy(vs, k, xs, want_caches=False)
        o.scores(vs, k, xs)
        n += 1
    dt = (time.perf_counter() - t0) / n
    per_step = dt * R / nr
    return {"value": 1.0 / per_step, "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": f"{n} x (energy + scores) of {nr} replica over {T} steps, scaled x{R}",
            "cpu_model": _cpu_model(), "nproc": os.cpu_count()}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
