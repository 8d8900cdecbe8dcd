#!/usr/bin/env python3
"""Throughput of the plain GPU replica sampler on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--replicas R] [--hours T] [--fp64]

A step is one replica step (``replica_step!``, src/replicas.jl:676-705, ``replica_inner = 1``)
over all replicas of one chain: every replica's weather step (pCN bursts staggered over the
replicas, regime windows, compensated re-dispatch), one operation sweep of exact line moves
(every atom, round(4N) random moves per step over the plain families at the reference's
weights, the storage multiscale sweep with probability 1/2 and the long-range storage sweep),
the energy totals and the capacity score, one Langevin step of the shared capacities
and the dispatch following on every replica. The workload is the 40-node / 70-line synthetic
system at an hourly horizon (8760 steps) with 128 weather replicas (the configuration of the
learned-proposal target); the state is FP32 with FP64 totals, the Jacobian test and the Langevin
step in FP64. Inputs are seeded and resident in HBM before the timed region.
Beside the timed steps: the FP64 oracle on the host cores from the device's own state (CPU
baseline), the learned proposal, the uneven batch, and the ESS/s of total system cost (replica
driver, plain against learned + batched) and of the operational cost at frozen capacities and
weather (plain sweeps against plain + learned sweeps).
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "esm-sampling_amd"))

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--replicas", type=int, default=128)
    ap.add_argument("--hours", type=int, default=8760)
    ap.add_argument("--nodes", type=int, default=40)
    ap.add_argument("--lines", type=int, default=70)
    ap.add_argument("--fp64", action="store_true", help="FP64 device state")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal leg")
    ap.add_argument("--no-ess", action="store_true",
                    help="skip the ESS legs (the replica driver with and without the learned proposal, ESS/s of "
                         "total system cost; the operational cost at frozen capacities and weather)")
    ap.add_argument("--ess-steps", type=int, default=48, help="replica steps of each driver run of the ESS leg "
                                                               "(1/6 warmup)")
    ap.add_argument("--ess-chains", type=int, default=2, help="chains of each driver run of the ESS leg")
    ap.add_argument("--ess-sweeps", type=int, default=40,
                    help="production sweeps of each run of the frozen-(k, x) dispatch ESS leg")
    ap.add_argument("--ess-batch", choices=("threads", "graph"), default="graph",
                    help="how the learned run of --ess runs its chains' sweeps (the plain run: threads)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="threads of the CPU baseline (0: all cores)")
    ap.add_argument("--p-storage-ms", type=float, default=0.5,
                    help="probability of the storage multiscale sweep per operation sweep (the reference's 1/2)")
    ap.add_argument("--xpair", type=float, default=0.1,
                    help="weight of the cross-node storage pair sxpair in the local sweep (storage_xpair, 0.1)")
    ap.add_argument("--xlong", type=int, default=8,
                    help="interval storage pairs sxlong per unit in the long-range sweep (storage_xlong, 8)")
    ap.add_argument("--tune", choices=("table", "live", "off"), default="table",
                    help="launch form of the random families: from the committed tuning table when it "
                         "holds this shape (default), tuned on the device now, or the library default")
    ap.add_argument("--no-batch", action="store_true", help="skip the uneven-batch leg")
    ap.add_argument("--copy-calib", action="store_true",
                    help="first run a plain device copy of 1 GiB three times (the calibration kernel of the "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counter runs)")
    return ap.parse_args()


class Workload:
    """The device-resident replica set and the controls of the replica step."""

    def __init__(self, N, L, T, R, prec, seed=1):
        import numpy as np
        from esmsampling import api, synth
        self.np, self.api = np, api
        # one weather year over T steps; longer horizons are several years at hourly resolution
        self.s = synth.make_system(N=N, L=L, T=T, dt=8760.0 / T if T <= 8760 else 1.0, seed=seed)
        self.R, self.seed = R, seed
        self.dev = api.Replicas(self.s, R, prec)
        api.sample_weather(self.dev, key=(seed, 7), counter=0)
        _, _, x = self.dev.download()
        v = synth.heuristic_state(self.s, x)
        self.dev.upload(v=v, k=self.s.k0)
        self.w = api.uniform_weights(np.ones(R, dtype=bool))
        self.eta = np.full(6, 5e-4)
        self.ks = np.array([1000.0, 500.0, 200.0, 300.0, 100.0, 100.0])   # κ_s per carrier (src/replicas.jl:522)
        self.rng = np.random.default_rng(seed + 100)
        self.burst_every, self.phi, self.window, self.window_prob = 40, 0.3, 60, 0.5

    def controls(self, it):
        np, api = self.np, self.api
        th = np.zeros(self.R)
        wp = np.zeros(self.R)
        for j in range(self.R):
            th[j], burst = api.weather_schedule(it, j, self.R, 0.0, self.burst_every, self.phi)
            wp[j] = self.window_prob if burst else 0.0
        return th, wp

    def step(self, it):
        api, dev = self.api, self.dev
        th, wp = self.controls(it)
        api.weather_process_step(dev, th, wp, window=self.window, compensate=True, key=(self.seed, 11), counter=it)
        self.last_stats = api.operation_sweep(dev, 1, key=(self.seed, 13), counter=it)
        parts = api.energy_cost(dev)
        sk, fk = api.capacity_scores(dev)
        xi = self.rng.standard_normal((6, self.s.N))
        knew, _ = api.langevin_capacities(dev, self.w, xi, self.eta, self.ks, kmin=1.0, clip=0.02, box=True)
        api.follow_capacities(dev, knew, kmin=1.0)
        return parts


def sweep_unit_bytes(s, es):
    """Algorithmic bytes of an operation sweep per (replica, hour): the v coordinates (4N − 1 of
    them carry information: one angle is the gauge), r and F read and written once, and the
    weather inputs (3 capacity factors and the load per node) read once: 2 (4N − 1 + N + L) + 4N
    values (2,792 B in FP32 at 40 nodes / 70 lines)."""
    return (2 * (4 * s.N - 1 + s.N + s.L) + 4 * s.N) * es


def kernel_bytes(s, R, es, ntheta_launch, stats, prof, steps):
    """Algorithmic HBM bytes per launch of each kernel (reads + writes of state and caches that
    the kernel's algorithm needs, each value once). For the random families: the moves of one
    launch (from the sweep statistics) times the values one move reads and writes."""
    import numpy as np
    N, L, T = s.N, s.L, s.T
    half = T // 2
    deg = 2.0 * L / N
    out = {
        "k_energy": (es * (7 * N + N + L) + 6 * 8) * T * R,        # θ, gas, 2 SoC, 3 CF in; r, F, 6 step parts out
        "k_gen": 4 * es * N * T * R,                                 # y, r in and out
        "k_theta": (8 * L + 4 * N) * es * T * R / max(ntheta_launch, 1),   # per site 4·deg + 4 values
        "k_soc": 8 * es * N * half * R,                              # e(τm, τ, τ+1), r(τm, τ) in; e(τ), r(τm, τ) out
        "k_follow": 9 * es * N * T * R,                              # y, r, 3 CF, 2 SoC in; y, r out
    }
    # per move: unit y, r; potential θ, deg flows, deg + 1 balances; spair 3 SoC, 2 gas, 2 r (+ writes);
    # cell y, r of every node of the cell (mean N/3); xloc/xlocc 2 gas, ~2·deg flows and balances; xpair
    # 2 gas, every flow and balance and angle
    per_move = {"k_unit": 4, "k_pot": 2 * (2 * deg + 2), "k_spair": 12, "k_cell": 4 * N / 3,
                "k_xloc": 2 * (2 + 4 * deg), "k_xlocc": 2 * (2 + 6 * deg), "k_xpair": 2 * (2 + L + 2 * N)}
    fam = {"k_spair": 3, "k_unit": 4, "k_pot": 5, "k_cell": 6, "k_xloc": 7, "k_xlocc": 8, "k_xpair": 9}
    for k, f in fam.items():
        launches = max(prof.get(k, (0, 0))[1] / max(steps, 1), 1)
        out[k] = per_move[k] * es * float(stats[f, 0]) / launches
    # the merged launch of the table-driven families (one per parity class) on the per-(replica, hour)
    # basis of a sweep: every operation coordinate (θ, gas, two SoC: 4N), the residual and flow
    # caches read and written once, the weather inputs (3 CF + load: 4N) read once, over the T/2
    # steps of the class. A per-move sum would count the values that several moves of a step touch
    # again for every move.
    merged = ("k_cell", "k_xloc", "k_xlocc", "k_xpair")
    out["k_random"] = sweep_unit_bytes(s, es) * (T // 2) * R
    for k in merged:
        if prof.get(k, (0, 0))[1] == 0:
            out.pop(k)
    # storage multiscale sweep (when it runs: 2 launches per block length L = 2, 4, …, 128 < T/2): a
    # move on a block of L + 1 steps reads ye, ye1 and r of every step and the gas at both ends of
    # each of its nu units, and writes r of every step, ye of the L inner instants and the gas at the
    # ends: (5L + 8)·nu values; nu = 1/2 + 1/2·(mean country size); 8 moves per block
    lengths = [L for L in (2, 4, 8, 16, 32, 64, 128) if L < T // 2]
    if lengths:
        nu = 0.5 + 0.5 * float(np.bincount(s.country).mean()) if s.ncountry > 0 else 1.0
        tot = sum(8 * (T - L) / L * (5 * L + 8) * nu * es * R for L in lengths)
        out["k_storage_ms"] = tot / (2 * len(lengths))
    # long-range storage sweep (one launch per step): a plateau reads the L + 1 instants of its unit's
    # state of charge, r and the gas at both end steps, and writes the L rising instants, r and the
    # gas at the ends: (2L + 9)·es with L = T/2 on average; a level move reads and writes T instants
    nm, nl = float(stats[11, 0]), float(stats[12, 0])
    if nm + nl > 0:
        out["k_storage_long"] = es * (nm * (2 * (T / 2) + 9) + nl * 2 * T)
    # storage pairs: sxpair (one launch per parity class) on the merged launch's per-(replica, hour)
    # basis; sxlong (one launch per long-range sweep): a move reads both units' instants and the
    # residuals of their nodes over its interval (T/2 steps on average) and writes the instants:
    # 6 values per interior step
    if float(stats[13, 0]) > 0:
        out["k_sxpair"] = sweep_unit_bytes(s, es) * (T // 2) * R
    nx = float(stats[14, 0])
    if nx > 0:
        out["k_sxlong"] = es * nx * 6 * (T / 2)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X (asked for --gpus {args.gpus}); the multi-GPU form (one RCCL "
              f"allreduce of the weighted score per step) is designed, not run", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from esmsampling import FP32, FP64, _lib, api

    prec = FP64 if args.fp64 else FP32
    es = 8 if prec == FP64 else 4
    if args.copy_calib:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")     # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    T, R = args.hours, args.replicas
    tw = time.perf_counter()
    wl = Workload(args.nodes, args.lines, T, R, prec)
    print(f"workload ready ({args.nodes} n x {T} h x {R} r): {time.perf_counter() - tw:.1f} s", file=sys.stderr,
          flush=True)
    api.sweep_options(wl.dev, p_storage_ms=args.p_storage_ms)
    api.storage_cross_options(wl.dev, p_xpair=args.xpair, nxlong=args.xlong)
    wl.xpair, wl.xlong = args.xpair, args.xlong
    launch = apply_tuning(api, wl.dev, args.tune)
    lib, h = wl.dev.lib, wl.dev.handle
    tw = time.perf_counter()
    for it in range(args.warmup):
        wl.step(it)
        print(f"warmup step {it + 1}/{args.warmup}: {time.perf_counter() - tw:.1f} s", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    lib.esm_profile(h, 1)
    t0 = time.perf_counter()
    tp = t0
    for it in range(args.warmup, args.warmup + args.steps):
        parts = wl.step(it)
        if time.perf_counter() - tp > 20.0:   # a progress line on long steps (the full-size configuration)
            tp = time.perf_counter()
            print(f"timed step {it + 1 - args.warmup}/{args.steps}: {tp - t0:.1f} s", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    names = {_lib.KERN_ENERGY: "k_energy", _lib.KERN_SCORE: "k_score", _lib.KERN_WEATHER: "weather",
             _lib.KERN_GEN: "k_gen", _lib.KERN_THETA: "k_theta", _lib.KERN_SOC: "k_soc", _lib.KERN_SPAIR: "k_spair",
             _lib.KERN_FOLLOW: "k_follow", _lib.KERN_SWEEP: "sweep", _lib.KERN_UNIT: "k_unit",
             _lib.KERN_POT: "k_pot", _lib.KERN_CELL: "k_cell", _lib.KERN_XLOC: "k_xloc", _lib.KERN_XLOCC: "k_xlocc",
             _lib.KERN_XPAIR: "k_xpair", _lib.KERN_SMS: "k_storage_ms", _lib.KERN_SLONG: "k_storage_long",
             _lib.KERN_RANDOM: "k_random", _lib.KERN_SXPAIR: "k_sxpair", _lib.KERN_SXLONG: "k_sxlong"}
    prof = {}
    for kid, name in names.items():
        tot, cnt = C.c_double(), C.c_int64()
        lib.esm_profile_read(h, kid, C.byref(tot), C.byref(cnt))
        prof[name] = (tot.value, cnt.value)
    lib.esm_profile(h, 0)
    ms_step = 1e3 * (t1 - t0) / args.steps
    value = args.steps / (t1 - t0)
    ntheta = prof["k_theta"][1] / max(args.steps, 1)
    kb = kernel_bytes(wl.s, R, es, ntheta, wl.last_stats, prof, args.steps)
    cand = {k: prof[k][0] for k in kb}
    dom = max(cand, key=cand.get)
    tot_ms, cnt = prof[dom]
    kern_ms = tot_ms / max(cnt, 1)
    achieved = kb[dom] / (kern_ms * 1e-3) / 1e9
    breakdown = {k: round(v[0] / args.steps, 4) for k, v in prof.items()}
    traffic = measured_traffic(dom, wl.s, T, R, prec, launch["form"])
    say(f"timed: {value:.3f} steps/s, {ms_step:.1f} ms per step; dominant kernel {dom} {kern_ms:.2f} ms")
    line_moves = int(wl.last_stats[:, 0].sum())   # exact line moves of one sweep (atoms and random families)
    out = {
        "metric": "sweeps/s (replica steps/s of the plain GPU replica sampler)",
        "value": value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == FP64 else "f32",
        "data": "synthetic (seeded 40-node/70-line system, weather drawn from the VAR prior on the device)",
        "config": {"workload": f"replica sampler step, {wl.s.N} nodes / {wl.s.L} lines, {T} hourly steps, "
                               f"{R} weather replicas (weather + operation sweep + energy + score + Langevin + "
                               f"following; storage pairs sxpair {args.xpair:g}, sxlong {args.xlong})",
                   "replicas": R, "hours": T, "nodes": wl.s.N, "lines": wl.s.L,
                   "line_moves_per_step": line_moves, "line_moves_per_s": line_moves * value,
                   "launch": launch},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel_ms": kern_ms,
                     "bytes_per_launch": kb[dom], "launches": cnt,
                     "basis": (f"{sweep_unit_bytes(wl.s, es)} B per (replica, hour) x T/2 hours of the parity "
                               f"class x R" if dom == "k_random" else "per-kernel algorithmic bytes"),
                     "sweep": {"bytes_per_sweep": sweep_unit_bytes(wl.s, es) * T * R,
                               "ms_per_sweep": prof["sweep"][0] / max(prof["sweep"][1], 1),
                               "achieved": sweep_unit_bytes(wl.s, es) * T * R /
                                           (1e-3 * prof["sweep"][0] / max(prof["sweep"][1], 1)) / 1e9}},
        "kernel_ms_per_step": breakdown,
        "sweep_stats": {name: {"moves": int(st[0]), "accepted": int(st[1]), "improper": int(st[2]),
                               "cut": int(st[3]), "evals_per_move": round(float(st[4]) / max(int(st[0]), 1), 3)}
                        for name, st in zip(api.FAMILY_NAMES, wl.last_stats)},
        "total_cost_mean": float(np.mean(parts[:, 0] + parts[:, 1])),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_threads, args.warmup + args.steps)
        say(f"cpu baseline: {out['cpu_baseline']['value']:.4g} steps/s")
    wl.dev.close()
    if not args.no_learned:
        out["learned"] = learned_leg(args, prec)
        say(f"learned leg: {out['learned']['learned_sweep_ms']:.1f} ms per learned sweep")
    if not args.no_batch:
        out["batch"] = batch_leg(prec)
    if not args.no_ess:
        out["ess"] = ess_leg(args)
        out["ess_dispatch"] = dispatch_ess_leg(args, prec)
    print(json.dumps(out), flush=True)


def say(msg):
    print(msg, file=sys.stderr, flush=True)


def ess_leg(args):
    """ESS/s of total system cost (``ess_bulk`` of the replica-mean total, src/convergence.jl:111,
    over the production steps per second of production wall time) of the replica driver, plain and
    with the learned proposal (training before the first step and in the warmup; its time counted
    in ``ess_per_s_with_training``), on ``--ess-chains`` chains of the bench's replica set (40 nodes,
    8760 h, 128 replicas by default). The plain run sweeps its chains from host threads; the
    learned run as ``--ess-batch`` says (by default one captured graph per step: the learned +
    batched sampler). R-hat is the rank-normalized split R-hat of the same traces."""
    from esmsampling import replicas as RP
    nw = max(args.ess_steps // 6, 2)
    base = dict(init="heuristic", nchains=args.ess_chains, replicas=args.replicas, mlda_warmup=nw,
                nsamples=args.ess_steps - nw,
                replica_init_sweeps=10, synthetic_nodes=args.nodes, synthetic_lines=args.lines,
                snapshots=args.hours, resolution_hours=8760.0 / args.hours, precision="fp32", seed=7)
    # a short run first (discarded): the first run of a process pays one-time costs (code objects
    # loaded at their first launch, clocks) that would otherwise fall on the plain run alone
    RP.run_replicas(RP.load_config({**base, "replicas": 16, "snapshots": 2920, "resolution_hours": 3.0,
                                    "mlda_warmup": 2, "nsamples": 4, "learned_proposal": True,
                                    "learned_train_steps": 10}))
    res = {}
    for name, extra in (("plain", dict(sweep_batch="threads")), ("learned", dict(learned_proposal=True,
                                                                                learned_train_steps=1000,
                                                         learned_train_every=10, learned_train_more=100,
                                                         sweep_batch=args.ess_batch))):
        r = RP.run_replicas(RP.load_config({**base, **extra}),
                            log=lambda m: print(m, file=sys.stderr, flush=True))
        tot = r["summary"].get("total", {})
        res[name] = {"ess_total": r["ess_total"], "ess_per_s": r["ess_per_s"],
                     "ess_per_s_with_training": r["ess_per_s_with_training"], "rhat": tot.get("rhat"),
                     "steps_per_s": r["steps_per_s"], "production_s": r["wall"]["production"],
                     "learned": r["learned"]}
        say(f"ess leg: {name} ESS {r['ess_total']} in {r['wall']['production']:.1f} s, R-hat {tot.get('rhat')}")
    p, l = res["plain"]["ess_per_s"], res["learned"]["ess_per_s_with_training"]
    res["ratio_learned_vs_plain"] = None if not p or l is None else l / p
    res["config"] = {"nchains": args.ess_chains, "replicas": args.replicas, "hours": args.hours,
                     "steps": args.ess_steps, "warmup": nw, "learned_sweep_batch": args.ess_batch}
    return res


def dispatch_ess_leg(args, prec, K=8, B=24, train_steps=1000, warm=20):
    """ESS/s of the operational cost at frozen capacities and weather: the dispatch observable that
    the learned proposal acts on. One replica set (the bench's shape); every replica is its own
    chain of its own target (its weather). After ``warm`` plain sweeps from the heuristic start,
    two runs from the same state: ``--ess-sweeps`` plain operation sweeps, and as many plain
    sweeps each followed by a learned sweep (the network trained first, its time counted). Per
    replica the bulk ESS of its op-cost trace (split chain, src/convergence.jl:111); the sum over
    the replicas per second of wall time, and the z-score of the two runs' replica-mean op cost
    (matched marginals: both runs sample the same target)."""
    import numpy as np
    import torch
    from esmsampling import api, synth
    from esmsampling.diagnostics import ess_bulk
    s = synth.make_system(N=args.nodes, L=args.lines, T=args.hours, dt=8760.0 / args.hours, seed=3)
    R = args.replicas
    dev = api.Replicas(s, R, prec)
    apply_tuning(api, dev, args.tune)
    api.sample_weather(dev, key=(8, 1))
    dev.upload(v=synth.heuristic_state(s, dev.download()[2]), k=s.k0)
    api.operation_sweep(dev, warm, key=(8, 2), counter=0)
    v0, _, _ = dev.download()
    res = {}
    for name in ("plain", "learned"):
        dev.upload(v=v0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        net = None
        if name == "learned":
            net = api.LearnedProposal(dev, block_hours=B, ncomp=K, key=(8, 3))
            net.train(train_steps, batch=4096, lr=2e-4, key=(8, 4))
        tr = np.empty((args.ess_sweeps, R))
        att = acc = 0
        for q in range(args.ess_sweeps):
            api.operation_sweep(dev, 1, key=(8, 5), counter=1000 + q)
            if net is not None:
                st = net.propose(key=(8, 6), counter=q)
                att += int(st[0])
                acc += int(st[1])
            tr[q] = api.energy_cost(dev)[:, 1]
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ess = np.array([ess_bulk([tr[:, j]]) for j in range(R)])
        res[name] = {"ess_sum": float(ess.sum()), "ess_median": float(np.median(ess)), "wall_s": wall,
                     "ess_per_s": float(ess.sum()) / wall, "op_mean": float(tr.mean()),
                     "op_se": float(np.sqrt(np.sum(tr.var(axis=0, ddof=1) / np.maximum(ess, 1.0))) / R),
                     "learned_acceptance": acc / att if att else None}
        say(f"dispatch ess: {name} ESS {ess.sum():.1f} over {R} replicas in {wall:.1f} s")
    dev.close()
    p, l = res["plain"], res["learned"]
    res["ratio_learned_vs_plain"] = l["ess_per_s"] / p["ess_per_s"] if p["ess_per_s"] > 0 else None
    res["z_op_mean"] = (l["op_mean"] - p["op_mean"]) / math.sqrt(l["op_se"] ** 2 + p["op_se"] ** 2)
    res["config"] = {"nodes": s.N, "hours": s.T, "replicas": R, "sweeps": args.ess_sweeps, "warm_sweeps": warm,
                     "learned": {"block_hours": B, "components": K, "train_steps": train_steps}}
    return res


MFMA_BF16_PEAK_TFLOPS = 2500.0   # MI355X dense BF16 MFMA


def learned_leg(args, prec, R=128, K=8, B=24, train_steps=300, sweeps=3):
    """The learned dispatch-block proposal on the 40-node / 8760 h / 128-replica workload: states
    of the plain sampler (two operation sweeps), maximum-likelihood training (Adam, batch 4096),
    then learned sweeps (every (node, 24 h block, replica) once: features, the MLP's forward pass
    on MFMA, draw, exact MH ratio, commit) beside one plain operation sweep of the same replicas.
    The forward pass's BF16 MFMA rate is its algorithmic flops (the network's shape) over the sum
    of its GEMM launches' durations (HIP events around every launch)."""
    import numpy as np
    import torch
    from esmsampling import _lib, api, synth

    s = synth.make_system(N=args.nodes, L=args.lines, T=args.hours, dt=1.0, seed=1)
    dev = api.Replicas(s, R, prec)
    api.sample_weather(dev, key=(5, 1))
    dev.upload(v=synth.heuristic_state(s, dev.download()[2]), k=s.k0)
    api.operation_sweep(dev, 1, key=(5, 2), counter=0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    api.operation_sweep(dev, 1, key=(5, 2), counter=1)
    plain_ms = 1e3 * (time.perf_counter() - t0)
    net = api.LearnedProposal(dev, block_hours=B, ncomp=K, key=(5, 3))
    lib, h = dev.lib, dev.handle
    lib.esm_profile(h, 1)
    t0 = time.perf_counter()
    loss = net.train(train_steps, batch=4096, lr=2e-4, key=(5, 4))
    train_ms = 1e3 * (time.perf_counter() - t0) / train_steps
    net.propose(key=(5, 5), counter=0)             # warm
    att = acc = 0
    t0 = time.perf_counter()
    for c in range(sweeps):
        st = net.propose(key=(5, 5), counter=1 + c)
        att += int(st[0])
        acc += int(st[1])
    sweep_ms = 1e3 * (time.perf_counter() - t0) / sweeps
    prof = {}
    for kid, name in ((_lib.KERN_MLPF, "forward"), (_lib.KERN_MLPP, "propose"), (_lib.KERN_MLPT, "train_step")):
        tot, cnt = C.c_double(), C.c_int64()
        lib.esm_profile_read(h, kid, C.byref(tot), C.byref(cnt))
        prof[name] = (tot.value, cnt.value)
    lib.esm_profile(h, 0)
    dev.close()
    rows = net.n_inst * (sweeps + 1)                # forward rows timed (the warm sweep included)
    flops_row = 2.0 * (net.din * net.dh + 2 * net.dh * net.dh + net.dh * net.dout)
    fwd_ms, launches = prof["forward"]              # the sum of the GEMM launches' own durations
    tflops = flops_row * rows / (fwd_ms * 1e-3) / 1e12
    return {"workload": f"{s.N} nodes, {s.T} h, {R} replicas, {B} h blocks, {K} components per hour",
            "blocks_per_sweep": net.n_inst, "learned_sweep_ms": sweep_ms, "plain_operation_sweep_ms": plain_ms,
            "acceptance": acc / max(att, 1), "train_steps": train_steps, "train_step_ms": train_ms,
            "train_loss_first_last": [float(loss[0]), float(loss[-1])],
            "forward_ms_per_sweep": fwd_ms / (sweeps + 1), "propose_kernel_ms_per_sweep": prof["propose"][0] / (sweeps + 1),
            "mfma": {"kernel": "k_mlp_fwd (fused forward pass, 4 layers)", "bound": "mfma", "achieved": tflops,
                     "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tflops / MFMA_BF16_PEAK_TFLOPS,
                     "flops_per_row": flops_row, "rows": rows, "launches": launches,
                     "kernel_ms_mean": fwd_ms / max(launches, 1)}}


def apply_tuning(api, dev, mode):
    """The launch form and workgroup size of the random families: from the committed tuning table
    (``esmsampling/tuning.json``, written by ``python -m esmsampling.tune`` on the GPU) when it
    holds this shape, tuned now on the device (``live``), or the library default."""
    from esmsampling import tune
    if mode == "table" and tune.TABLE.exists():
        table = json.loads(tune.TABLE.read_text())
        if api._shape_key(dev) in table:
            r = api.autotune(dev, cache=str(tune.TABLE))
            return {"form": r["form"], "block": r["block"], "pad": r["pad"], "source": "tuning table"}
    if mode == "live":
        r = api.autotune(dev, nsweeps=1)
        return {"form": r["form"], "block": r["block"], "pad": r["pad"], "source": "tuned now",
                "ms_per_sweep": {f"{f}/{b}/{p}": round(t, 2) for (f, b, p), t in r["ms"].items()}}
    return {"form": 1, "block": 256, "pad": 1, "source": "library default"}


def batch_leg(prec, reps=3):
    """Uneven batch (40 and 256 nodes, 2920 and 8760 h, 256 replicas in all): one operation sweep
    of every replica set as one captured HIP graph (``api.Batch``) against the same sweeps run
    one set after the other, each set at its tuned launch form, with both storage sweeps at the
    reference's probabilities (multiscale 1/2, long-range 1). The graph is instantiated once and
    updated in place on every later call, however the storage sweeps' draws come out."""
    from esmsampling import api, synth
    shapes = [(40, 70, 2920, 96), (40, 70, 8760, 64), (256, 450, 2920, 64), (256, 450, 2920, 32)]
    sets = []
    for i, (N, L, T, R) in enumerate(shapes):
        s = synth.make_system(N=N, L=L, T=T, dt=1.0, seed=1 + i)
        d = api.Replicas(s, R, prec)
        api.sample_weather(d, key=(1, i))
        d.upload(v=synth.heuristic_state(s, d.download()[2]), k=s.k0)
        api.sweep_options(d, p_storage_ms=0.5)
        apply_tuning(api, d, "table")
        sets.append(d)
    keys = [(2, i) for i in range(len(sets))]
    b = api.Batch(sets)

    def one_by_one(c):
        for i, d in enumerate(sets):
            api.operation_sweep(d, 1, key=keys[i], counter=c)

    def graph(c):
        b.sweeps(1, keys=keys, counters=[c] * len(sets))

    res = {}
    for k, (name, f) in enumerate((("one_by_one", one_by_one), ("graph", graph))):
        f(100 * k)
        t0 = time.perf_counter()
        for c in range(reps):
            f(100 * k + 1 + c)
        res[name] = 1e3 * (time.perf_counter() - t0) / reps
        print(f"batch leg: {name} {res[name]:.1f} ms per batched sweep", file=sys.stderr, flush=True)
    counts = b.counts()
    b.close()
    for d in sets:
        d.close()
    return {"workload": "one operation sweep of 4 replica sets: 40 n x 2920 h x 96 r, 40 n x 8760 h x 64 r, "
                        "256 n x 2920 h x 64 r, 256 n x 2920 h x 32 r (storage multiscale sweep at p = 1/2, "
                        "long-range storage sweep on)",
            "ms_one_by_one": res["one_by_one"], "ms_graph": res["graph"],
            "speedup": res["one_by_one"] / res["graph"], "graph_launches_instantiations_updates": counts}


def traffic_profile(s, T, R, prec):
    """The committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload, if there is one."""
    from esmsampling import FP32
    p = ROOT / "profiles" / f"hbm_traffic_replica_step_{s.N}n_{T}h_{R}r_{'fp32' if prec == FP32 else 'fp64'}.json"
    return p if p.exists() else None


def measured_traffic(kernel, s, T, R, prec, form=1):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (calibrated against a plain copy), when it was measured on this exact workload."""
    p = traffic_profile(s, T, R, prec)
    if p is None:
        return None
    names = {"k_spair": "k_rand_spair<float, true>", "k_soc": "k_soc<float>", "k_gen": "k_gen<float>",
             "k_energy": "k_energy<float, true, false, true>", "k_follow": "k_follow<float>",
             "k_unit": "k_rand_unit<float, true>", "k_cell": "k_sparse<float, 0, true>",
             "k_xloc": "k_sparse<float, 1, true>", "k_xlocc": "k_sparse<float, 2, true>",
             "k_xpair": "k_sparse<float, 3, true>", "k_storage_ms": "k_storage_ms",
             "k_storage_long": "k_storage_long<float, true>",
             "k_random": f"k_sparse_all<float, true, {'true' if form > 1 else 'false'}>",
             "k_sxlong": "k_sxlong<float, 1>"}
    got = json.loads(p.read_text())["bytes_per_dispatch"]
    if kernel == "k_sxpair":   # the pair kernel's form (mixed 16/32 lanes, or 64 lanes) is in its name
        for name in ("k_rand_sxpair_mix<float>", "k_rand_sxpair<float, 64, 1>", "k_rand_sxpair<float, 64, 2>"):
            if name in got:
                return got[name]
        return None
    return got.get(names.get(kernel, ""))


def cpu_baseline(wl, threads, it):
    """The FP64 C oracle on the host cores: one full replica step (weather step, operation sweep,
    energy, score, Langevin, following) of a bounded sample of the replicas, started from the
    device's own state after the timed steps (the same operation, weather, regimes and
    capacities, so the walks are as long as the GPU's), scaled to all replicas."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np
    import oracle
    from esmsampling import api
    nth = threads or os.cpu_count() or 1
    s, R = wl.s, wl.R
    rs = max(1, min(R, nth))
    o = oracle.Oracle(s)
    v, k, x = wl.dev.download()
    m, z = api.weather_state(wl.dev)
    v, x, m, z = (np.ascontiguousarray(a[..., :rs]) for a in (v, x, m, z))
    th, wp = wl.controls(it)
    t0 = time.perf_counter()
    m, z, x, v, _ = o.weather_step(m, z, x, v, k, th[:rs], wp[:rs], window=60, compensate=True, key=(1, 11),
                                   counter=it)
    v, r, F, _ = o.sweeps(v, k, 1, key=(1, 13), counter=it, x=x, threads=nth, p_xpair=wl.xpair, xlong_moves=wl.xlong)
    o.energy(v, k, x, want_caches=False)
    sk, fk, _ = o.scores(v, k, x)
    knew, _ = o.langevin(k, sk, fk, np.full(rs, 1.0 / rs), np.zeros((6, s.N)), wl.eta, wl.ks)
    o.follow(v, k, x, knew)
    dt = time.perf_counter() - t0
    per_step = dt * R / rs
    return {"value": 1.0 / per_step, "unit": "steps/s", "cores": nth, "kind": "port",
            "sample": f"one full replica step of {rs} of the {R} replicas x {s.T} steps from the device state after "
                      f"the timed steps (oracle, OpenMP over replicas in the sweep), {dt:.1f} s, scaled x{R / rs:g}",
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
