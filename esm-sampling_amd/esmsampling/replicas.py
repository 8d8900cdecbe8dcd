"""The replica–Langevin driver (``sampler = "replicas"``) over the device path, on one GPU.

Mirrors ``run_replicas`` (src/replicas.jl:841-947) and its pieces: ``ReplicaCtl``
(:481-515), the gate (``relaxation_indicator`` :634-641, ``update_gate!`` :651-661), the drift
weights (:303-314, :664), ``replica_step!`` (:676-705), the lag probe and the step adaptation
(:717-773), ``init_replicas`` (:783-795) and the replica-mean monitor (:827) with the
convergence report of the drivers (src/chains.jl:301-310). One chain is one ``api.Replicas``
handle (its replicas share the capacities); every device call works on all replicas of the
chain at once.

Configuration: TOML with every key of the reference's ``Config`` (src/run.jl:7-127) and its
defaults; an unknown key raises, as ``load_config`` does (src/run.jl:135-156). Keys that the
replica path uses are implemented, or accepted at the values this package supports (others are
refused); keys that the replica path never reads are accepted and have no effect. The system is
the seeded synthetic one (``synth.make_system``): its nodes and hours per step come from the
finest level of ``levels`` (or the keys ``synthetic_nodes``, ``resolution_hours``), its regimes
from ``K``, its countries from the coarsest level's cells; ``synthetic_lines``, ``snapshots``,
``system_seed`` and ``precision`` have no counterpart in the reference (it reads its data file),
nor do the ``learned_*`` keys of the learned dispatch-block proposal (off by default).

Random streams: Philox4x32-10 keys per (seed, chain, purpose) replace the MersenneTwister streams
``restart_seed(cfg, r, 1000c + j)`` (src/replicas.jl:872-874); the Langevin normals come from a
numpy Philox generator per chain.
"""
from __future__ import annotations

import dataclasses
import functools
import math
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import FP32, FP64, api, records as rec_io, synth
from .diagnostics import convergence_summary, warmup_end

NK = synth.NK

# κ_s offsets of w = log(k + κ_s) per carrier (MW), src/replicas.jl:518
LANGEVIN_KAPPA = {"onwind": 500.0, "solar": 1000.0, "offwind": 200.0, "gas": 300.0, "battery": 100.0, "H2": 100.0}


# Default lines of a synthetic network of N nodes (the benchmark networks: 40 nodes / 70 lines and
# 256 nodes / 450 lines; otherwise 1.75 lines per node, at least a spanning tree).
def default_lines(N: int) -> int:
    return {40: 70, 256: 450}.get(N, max(N - 1, int(round(1.75 * N))))


@dataclasses.dataclass
class Config:
    """Every key of the reference's ``Config`` (src/run.jl:7-127) with its default, plus this
    driver's own keys (the synthetic system, the learned proposal, the launch form and the sweep
    batching). ``check`` sorts the reference keys into three kinds: those the replica path uses
    (implemented here), those it uses at values this package supports (others are refused with the
    reference's message style) and those the replica path never reads (accepted, no effect:
    level-0 warmup, bias learning, the chains' fine step, the device driver's capacity moves)."""
    # -- data and outputs (src/run.jl:8-14) --
    data: str = "data/processed/model.h5"
    out: str = ""
    cache_dir: str = ""
    cache_mirror: str = ""
    checkpoint: str = ""
    checkpoint_minutes: float = 30.0
    stop_file: str = ""
    # -- model and levels --
    levels: list = dataclasses.field(default_factory=lambda: [(10, 24), (40, 24)])
    seed: int = 1
    K: int = 4
    c_shed: float = 3000.0
    c_dump: float = 5.0
    c_viol: float = 3000.0
    local_extra: float = 4.0
    fine_kernel: str = "auto"
    cell_level_power: float = 1.0
    xloc_partner_hops: int = 1
    xloc_ball_hops: int = 0
    kop: bool = False
    storage_long: float = 1.0
    storage_long_moves: int = 32
    storage_long_hours: float = 24.0
    storage_xpair: float = 0.1
    storage_xlong: int = 8
    storage_xpair_hops: int = 1
    storage_xpair_ball: int = 1
    measure: str = "utilization"
    storage: str = "soc"
    renewables: str = "load"
    flow_theta: str = "nested"
    co2_price: float = math.nan
    co2_cap: float = 0.0
    brownfield: bool = False
    losses: bool = False
    loss_segments: int = 4
    thermal: str = "hard"
    thermal_band: float = 0.1
    link_losses: bool = False
    uc: list = dataclasses.field(default_factory=list)
    fine_weather: float = 0.0
    fine_weather_blocks: float = 1.0
    fine_weather_block_steps: int = 600
    init: str = "lp"
    basis: str = "haar"
    use_bias: bool = True
    beta_v: float = 4e-5
    beta_k: float = 4e-8
    calibrate_beta: bool = False
    eps_v: float = 0.01
    eps_k: float = 0.05
    warmup0: int = 1500
    bias_start: int = 150
    bias_every: int = 25
    bias_samples: int = 16
    bias_sample_steps: int = 4096
    bias_table_res: int = 100
    bias_warmstart: float = 20.0
    descent_iters: int = 4000
    mlda_warmup: int = 100
    nsamples: int = 300
    bias_monotone: bool = True
    bias_online: bool = True
    bias_online_every: int = 1
    bias_online_min: int = 16
    bias_trust: float = 1.0
    bias_post: bool = True
    # -- chains and convergence diagnostics --
    nchains: int = 4
    chain_starts: list = dataclasses.field(default_factory=lambda: ["low", "high", "heuristic", "heuristic"])
    rhat_max: float = 1.01
    ess_min: float = 400.0
    mcse_max: float = 0.05
    rhat_screen: float = 1.05
    # -- the replica–Langevin sampler --
    sampler: str = "chains"
    replicas: int = 8
    replica_init_sweeps: int = 40
    replica_inner: int = 1
    replica_theta: float = 0.0
    replica_burst_every: int = 40
    replica_burst_angle: float = 0.3
    replica_window: int = 60
    replica_window_prob: float = 0.5
    replica_ffbs: bool = False
    replica_compensate: bool = True
    replica_gate_min: int = 10
    replica_gate_max: int = 30
    replica_gate_min_window: int = 20
    replica_gate_max_window: int = 60
    replica_gate_win: int = 8
    replica_gate_tol: float = 0.3
    replica_probe_every: int = 20
    replica_probe_len: int = 30
    replica_probe_tail: int = 15
    replica_probe_win: int = 10
    replica_probe_avg: int = 3
    langevin_eta: float = 5e-4
    langevin_eta_min: float = 2e-5
    langevin_eta_max: float = 2e-3
    langevin_adapt: bool = True
    langevin_gap_hi: float = 0.10
    langevin_gap_lo: float = 0.05
    langevin_kappa: dict = dataclasses.field(default_factory=dict)
    langevin_kappa_default: float = 100.0
    langevin_kmin: float = 1.0
    langevin_clip: float = 0.02
    langevin_box: bool = True
    tilt_functional: str = "resload_mean"
    tilt_lambda: float = 0.0
    tilt_samples: int = 200
    # -- the device driver (device/): its batch sizes and capacity moves --
    device_backend: str = "gpu"
    device_sweeps: int = 10
    kmove_nprop: int = 2
    kmove_sigma: float = 0.3
    kpeak_sigma: float = 0.05
    device_sampler: str = "mh"
    verbose: bool = True
    # ---- keys of this driver (no counterpart in the reference) ----
    maxnats: float = 40.0
    # the learned dispatch-block proposal: after every operation sweep, `learned_moves` learned
    # moves of every (node, block, replica), exactly MH corrected; the network is trained by
    # maximum likelihood on the chain's own states, `learned_train_steps` Adam steps before the
    # first replica step and `learned_train_more` every `learned_train_every` warmup steps (never
    # in the production window, whose kernel stays fixed)
    learned_proposal: bool = False
    learned_block_hours: int = 24
    learned_components: int = 8
    learned_moves: int = 1
    learned_train_steps: int = 1000
    learned_train_every: int = 0
    learned_train_more: int = 200
    learned_batch: int = 4096
    learned_lr: float = 2e-4
    # the synthetic system that stands in for the data file: nodes and hours per step default to
    # the finest level (levels[-1] = (nodes, hours)), the horizon to one year, the lines to
    # default_lines(nodes)
    synthetic_nodes: int | None = None
    synthetic_lines: int | None = None
    snapshots: int | None = None
    resolution_hours: float | None = None
    system_seed: int = 1
    precision: str = "fp32"
    # launch form of the random families: "table" takes the committed tuning table's choice for
    # the replica sets' shape (esmsampling/tuning.json), "off" the library default
    launch_tuning: str = "table"
    # how the chains' operation sweeps run: "graph", the sweeps of all chains of a step as one
    # captured HIP graph (api.Batch), or "threads", each chain's sweep from its own host thread;
    # the states are the same
    sweep_batch: str = "graph"

    # the reference keys that the replica path never reads (src/replicas.jl:841-955 and the
    # kernels it builds): accepted with any value, no effect
    UNREAD = ("cache_dir", "cache_mirror", "fine_kernel", "kop", "flow_theta", "loss_segments", "thermal_band",
              "fine_weather", "fine_weather_blocks", "fine_weather_block_steps", "eps_v", "eps_k", "warmup0",
              "bias_start", "bias_every", "bias_samples", "bias_sample_steps", "bias_table_res", "bias_warmstart",
              "descent_iters", "bias_monotone", "bias_online", "bias_online_every", "bias_online_min", "bias_trust",
              "bias_post", "device_backend", "device_sweeps", "kmove_nprop", "kmove_sigma", "kpeak_sigma",
              "tilt_samples")

    @property
    def nodes(self) -> int:
        return int(self.synthetic_nodes if self.synthetic_nodes is not None else self.levels[-1][0])

    @property
    def lines(self) -> int:
        return int(self.synthetic_lines if self.synthetic_lines is not None else default_lines(self.nodes))

    @property
    def hours(self) -> float:
        return float(self.resolution_hours if self.resolution_hours is not None else self.levels[-1][1])

    @property
    def steps(self) -> int:
        return int(self.snapshots if self.snapshots is not None else round(8760 / self.hours))

    def check(self):
        """``check_config`` (src/run.jl:158-169) and the replica path's limits."""
        def refuse(key, msg):
            raise ValueError(f"{key} = {getattr(self, key)!r}: {msg}")
        # choice keys of the reference (its own messages)
        if self.storage != "soc":
            refuse("storage", 'only the state-of-charge form "soc" exists')
        if self.renewables != "load":
            refuse("renewables", 'only renewables as load "load" exist')
        if self.basis not in ("haar", "flow"):
            refuse("basis", '"haar" or "flow"')
        if self.init not in ("lp", "heuristic"):
            refuse("init", '"lp" or "heuristic"')
        if self.measure not in ("utilization", "physical"):
            refuse("measure", '"utilization" or "physical"')
        if self.fine_kernel not in ("auto", "local", "flow", "adaptive"):
            refuse("fine_kernel", '"auto", "local", "flow" or "adaptive"')
        if self.sampler not in ("chains", "replicas"):
            refuse("sampler", '"chains" or "replicas"')
        if self.device_sampler not in ("mh", "replicas"):
            refuse("device_sampler", '"mh" or "replicas"')
        if self.tilt_functional not in ("none", "resload_mean"):
            refuse("tilt_functional", '"none" or "resload_mean"')
        if self.tilt_lambda != 0 and self.tilt_functional == "none":
            raise ValueError("tilt_lambda ≠ 0 needs a tilt_functional")
        # values the replica path uses that this package does not implement
        if self.basis != "haar":
            refuse("basis", 'the flow basis and its transport families are not built here; use "haar"')
        if self.init != "heuristic":
            refuse("init", 'the reference LP start is not built here (no LP solver); use "heuristic"')
        if self.data != Config.data:
            refuse("data", "this driver runs the seeded synthetic system (synthetic_* keys); no data file is read")
        if self.calibrate_beta:
            refuse("calibrate_beta", "β from the LP needs the LP start; give beta_v and beta_k")
        if self.co2_cap > 0:
            refuse("co2_cap", "tuning the CO2 price to a cap is not built here (0: off)")
        if self.brownfield:
            refuse("brownfield", "brownfield floors and fleets are not built here")
        if self.losses:
            refuse("losses", "line losses are not built here (hard limits without losses)")
        if self.thermal != "hard":
            refuse("thermal", 'only hard line limits "hard" are built here')
        if self.link_losses:
            refuse("link_losses", "the synthetic system has no DC links")
        if self.uc:
            refuse("uc", "unit commitment is not built here")
        if self.cell_level_power != 1.0:
            refuse("cell_level_power", "the cell level is drawn ∝ its number of cells (1.0) here")
        if self.xloc_partner_hops != 1 or self.xloc_ball_hops != 0:
            raise ValueError("xloc_partner_hops = 1 and xloc_ball_hops = 0 (partners are neighbours, the flow is "
                             "re-routed on the pair) are the localized pair moves built here")
        if self.storage_xpair_hops != 1 or self.storage_xpair_ball != 1:
            raise ValueError("storage_xpair_hops = 1 and storage_xpair_ball = 1 are the storage pairs built here")
        if self.tilt_lambda != 0.0:
            refuse("tilt_lambda", "the stress tilt functional is not built (λ = 0 is the untilted drift)")
        if not (self.storage_xpair >= 0.0) or not (0 <= self.storage_xlong <= 256):
            raise ValueError("storage_xpair >= 0 and storage_xlong in [0, 256]")
        # levels: (cells, hours per step), coarse to fine (check_levels, src/run.jl:173-183); the
        # replica sampler runs at the finest level
        try:
            self.levels = [(int(a), int(b)) for a, b in self.levels]
        except (TypeError, ValueError):
            raise ValueError(f"levels = {self.levels!r}: a list of [cells, hours] pairs") from None
        if not self.levels:
            raise ValueError("levels: at least one level")
        for (M0, h0), (M1, h1) in zip(self.levels, self.levels[1:]):
            if M1 < M0 or h1 > h0 or h0 % h1 != 0:
                raise ValueError(f"levels {self.levels}: cells must not decrease and each level's hours must divide "
                                 f"the coarser level's")
        if self.hours <= 0 or self.steps < 2 or self.nodes < 2 or self.lines < self.nodes - 1:
            raise ValueError(f"synthetic system: {self.nodes} nodes, {self.lines} lines, {self.steps} steps of "
                             f"{self.hours} h")
        if not 2 <= self.K <= 4:
            refuse("K", "2 to 4 weather regimes are built here")
        if not self.local_extra >= 0:
            refuse("local_extra", "must be >= 0")
        if not (0.0 <= self.storage_long <= 1.0) or not (0 <= self.storage_long_moves <= 256) or self.storage_long_hours < 0:
            raise ValueError("storage_long in [0, 1], storage_long_moves in [0, 256], storage_long_hours >= 0")
        if self.learned_proposal and not (1 <= self.learned_block_hours <= min(24, self.steps)
                                          and 1 <= self.learned_components <= 8 and self.learned_moves >= 0
                                          and self.learned_train_steps >= 0 and self.learned_train_every >= 0
                                          and self.learned_train_more >= 0 and 1 <= self.learned_batch <= 65536
                                          and self.learned_lr > 0):
            raise ValueError("learned_*: block hours in [1, min(24, snapshots)], components in [1, 8], batch in "
                             "[1, 65536], lr > 0, counts >= 0")
        if self.sweep_batch not in ("threads", "graph"):
            raise ValueError("sweep_batch must be 'threads' or 'graph'")
        if self.launch_tuning not in ("table", "off"):
            raise ValueError("launch_tuning must be 'table' or 'off'")
        if self.precision not in ("fp32", "fp64"):
            raise ValueError('precision must be "fp32" or "fp64"')
        if self.replicas < 1 or self.nchains < 1:
            raise ValueError("replicas and nchains must be positive")
        for s in self.chain_starts:
            if s not in ("low", "high", "heuristic"):
                raise ValueError(f'chain start "{s}": "low", "high" or "heuristic"')
        return self

    def notes(self):
        """What this configuration leaves aside (logged at the start of a run, as the reference
        logs its ignored keys)."""
        out = []
        if self.use_bias:
            out.append("sampler = replicas: the learned bias is not used (use_bias ignored)")
        if self.data == Config.data:
            out.append(f"synthetic system: {self.nodes} nodes / {self.lines} lines, {self.steps} steps of "
                       f"{self.hours:g} h (seed {self.system_seed}) in place of {self.data}")
        return out


# Keys of old configuration files with the only value that matches the code (src/run.jl:133)
REMOVED_KEYS = {"kernel": "local", "p_fine": 1.0, "alpha0": False, "regularized": True}


def load_config(path_or_dict) -> Config:
    """``load_config`` (src/run.jl:135-156): a Config from a TOML file (or a dict of the same
    keys); an unknown key raises, a removed key must hold its one value, then ``check``."""
    if isinstance(path_or_dict, dict):
        d = dict(path_or_dict)
    else:
        import tomli
        with open(path_or_dict, "rb") as f:
            d = tomli.load(f)
    names = {f.name for f in dataclasses.fields(Config)}
    for k in list(d):
        if k in REMOVED_KEYS:
            if d[k] != REMOVED_KEYS[k]:
                raise ValueError(f"config key {k} = {d[k]} is no longer supported (the code always uses "
                                 f"{REMOVED_KEYS[k]})")
            del d[k]
        elif k not in names:
            raise ValueError(f"unknown config key {k}")
    return Config(**d).check()


@dataclasses.dataclass
class ReplicaCtl:
    """Settings of the replica sampler (``ReplicaCtl(cfg, P)``, src/replicas.jl:524-538)."""
    M: int
    inner: int
    theta: float
    burst_every: int
    phi: float
    window: int
    window_prob: float
    comp: bool
    gate_min: int
    gate_max: int
    gate_min_window: int
    gate_max_window: int
    gate_win: int
    gate_tol: float
    eta_w: np.ndarray
    kappa_s: np.ndarray
    kmin: float
    clip: float
    box: bool
    probe_every: int
    probe_len: int
    probe_tail: int
    probe_win: int
    probe_avg: int
    adapt: bool
    gap_hi: float
    gap_lo: float
    eta_min: float
    eta_max: float
    plong: float = 1.0
    long_moves: int = 32
    long_hours: float = 24.0
    xpair: float = 0.1
    xlong: int = 8
    tuning: str = "table"
    extra: float = 4.0
    measure: str = "utilization"
    ffbs: bool = False
    shrink: float = 0.7
    grow: float = 1.2

    @classmethod
    def from_config(cls, cfg: Config):
        ks = np.array([cfg.langevin_kappa.get(c, LANGEVIN_KAPPA.get(c, cfg.langevin_kappa_default))
                       for c in synth.CARRIERS], dtype=np.float64)
        return cls(M=cfg.replicas, inner=cfg.replica_inner, theta=cfg.replica_theta, burst_every=cfg.replica_burst_every,
                   phi=cfg.replica_burst_angle, window=cfg.replica_window, window_prob=cfg.replica_window_prob,
                   comp=cfg.replica_compensate, gate_min=cfg.replica_gate_min, gate_max=cfg.replica_gate_max,
                   gate_min_window=cfg.replica_gate_min_window, gate_max_window=cfg.replica_gate_max_window,
                   gate_win=cfg.replica_gate_win, gate_tol=cfg.replica_gate_tol,
                   eta_w=np.full(NK, cfg.langevin_eta), kappa_s=ks, kmin=cfg.langevin_kmin, clip=cfg.langevin_clip,
                   box=cfg.langevin_box, probe_every=cfg.replica_probe_every, probe_len=cfg.replica_probe_len,
                   probe_tail=cfg.replica_probe_tail, probe_win=cfg.replica_probe_win, probe_avg=cfg.replica_probe_avg,
                   adapt=cfg.langevin_adapt, gap_hi=cfg.langevin_gap_hi, gap_lo=cfg.langevin_gap_lo,
                   eta_min=cfg.langevin_eta_min, eta_max=cfg.langevin_eta_max, plong=cfg.storage_long,
                   long_moves=cfg.storage_long_moves, long_hours=cfg.storage_long_hours, xpair=cfg.storage_xpair,
                   xlong=cfg.storage_xlong, tuning=cfg.launch_tuning, extra=cfg.local_extra, measure=cfg.measure, ffbs=cfg.replica_ffbs)

    def sweep_options(self, dev):
        """The operation sweep's options of a replica set (``local_extra``, the long-range storage
        sweep's keys, the cross-node storage pairs ``storage_xpair`` / ``storage_xlong``), the
        reference measure of its scores and its launch form (from the tuning table when it holds
        the set's shape)."""
        api.sweep_options(dev, extra=self.extra)
        api.storage_long_options(dev, p=self.plong, nmoves=self.long_moves, hmin=self.long_hours)
        api.storage_cross_options(dev, p_xpair=self.xpair, nxlong=self.xlong)
        api.score_measure(dev, self.measure)
        if self.tuning == "table":
            from . import tune
            table = _tuning_table(str(tune.TABLE))
            if api._shape_key(dev) in table:
                api.autotune(dev, cache=str(tune.TABLE))


@functools.lru_cache(maxsize=4)
def _tuning_table(path: str) -> dict:
    import json
    import os
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def interior_nodes(k, khi, lo: float = 50.0):
    """Per k field the nodes with lo < k < box − 1 (src/replicas.jl:475)."""
    return (k > lo) & (k < khi - 1.0)


def masked_means(v, mask):
    """Mean of every field's node vector over its masked nodes, NaN where none (src/replicas.jl:477)."""
    out = np.full(v.shape[0], np.nan)
    for f in range(v.shape[0]):
        if mask[f].any():
            out[f] = v[f][mask[f]].mean()
    return out


def relaxation_indicator(gbuf, win: int) -> float:
    """Largest difference over the k fields between the mean normalised carrier score of the last
    win/2 steps and of the win/2 before; inf until there are ``win`` steps (src/replicas.jl:634-641)."""
    h = win // 2
    if not (len(gbuf) >= 2 * h and h >= 1):
        return math.inf
    n = len(gbuf)
    a = np.mean(gbuf[n - 2 * h:n - h], axis=0)
    b = np.mean(gbuf[n - h:n], axis=0)
    d = [abs(x - y) for x, y in zip(a, b) if math.isfinite(x) and math.isfinite(y)]
    return max(d) if d else 0.0


class ReplicaSet:
    """One chain's replicas (a device handle) and their gate, probe windows and Langevin step
    (``ReplicaSet``, src/replicas.jl:547-562)."""

    def __init__(self, dev: api.Replicas | None, eta_w, M: int | None = None):
        M = dev.R if dev is not None else int(M)
        self.dev = dev
        self.k = None
        self.since = np.full(M, 10 ** 9, dtype=np.int64)
        self.window = np.zeros(M, dtype=bool)
        self.active = np.ones(M, dtype=bool)
        self.gbuf = [[] for _ in range(M)]
        self.hist = [[] for _ in range(M)]
        self.eta_w = np.array(eta_w, dtype=np.float64)
        self.gaps = []


def update_gate(R: ReplicaSet, j: int, g, ctl: ReplicaCtl) -> bool:
    """``update_gate!`` (src/replicas.jl:651-661)."""
    if ctl.burst_every == 0:
        R.active[j] = True
        return True
    if R.active[j]:
        return True
    R.gbuf[j].append(np.asarray(g, dtype=np.float64))
    if len(R.gbuf[j]) > ctl.gate_win:
        R.gbuf[j].pop(0)
    gmin, gmax = (ctl.gate_min_window, ctl.gate_max_window) if R.window[j] else (ctl.gate_min, ctl.gate_max)
    if R.since[j] < gmin:
        return False
    R.active[j] = bool(R.since[j] >= gmax or relaxation_indicator(R.gbuf[j], ctl.gate_win) < ctl.gate_tol)
    if R.active[j]:
        R.gbuf[j].clear()
    return bool(R.active[j])


def drift_weights(active):
    """Drift weights without a tilt functional: 1/M_active over the active replicas (:664, :310-314)."""
    return api.uniform_weights(active)


class Keys:
    """Philox keys of one chain: (seed, 1000·chain + purpose)."""

    def __init__(self, seed: int, chain: int):
        base = 1000 * (chain + 1)
        self.sweep = (seed, base + 1)
        self.weather = (seed, base + 2)
        self.init = (seed, base + 3)
        self.learned = (seed, base + 4)
        self.train = (seed, base + 5)
        self.net = (seed, base + 6)
        self.probe = (seed, base + 500)


class Learned:
    """The learned proposal of one chain's replica set: the network, its schedule and counts."""

    def __init__(self, dev: api.Replicas, cfg: Config, keys: Keys):
        self.net = api.LearnedProposal(dev, block_hours=cfg.learned_block_hours, ncomp=cfg.learned_components,
                                       key=keys.net)
        self.cfg, self.keys = cfg, keys
        self.moves = cfg.learned_moves
        self.att = self.acc = 0
        self.ntrain = 0
        self.loss = []
        self.wall = 0.0          # seconds spent training

    def train(self, steps: int):
        if steps > 0:
            t0 = time.perf_counter()
            self.loss.extend(self.net.train(steps, batch=self.cfg.learned_batch, lr=self.cfg.learned_lr,
                                            key=self.keys.train, counter=self.ntrain))
            self.ntrain += 1
            self.wall += time.perf_counter() - t0

    def sweep(self, counter: int):
        """``moves`` learned moves of every (node, block, replica); returns (attempted, accepted)."""
        att = acc = 0
        for q in range(self.moves):
            st = self.net.propose(key=self.keys.learned, counter=counter * max(self.moves, 1) + q)
            att += int(st[0])
            acc += int(st[1])
        self.att += att
        self.acc += acc
        return att, acc


def _sweep_scores(R: ReplicaSet, nsweeps: int, key, counter: int, maxnats: float, r: float, learned=None,
                  sweeper=None):
    """``nsweeps`` operation sweeps (each followed by the learned moves, when there are), after
    each the energy (caches rebuilt) and the capacity scores; returns (mean of r·ŝ over the sweeps
    [R, 6, N], fk, parts of the last sweep, stats, learned (attempted, accepted)). ``sweeper(dev,
    key, counter, maxnats)`` runs one sweep and returns its statistics (``api.operation_sweep``
    when None)."""
    acc, fk, parts, stats = None, None, None, np.zeros((len(api.FAMILY_NAMES), 5), dtype=np.int64)
    lst = np.zeros(2, dtype=np.int64)
    for q in range(nsweeps):
        if sweeper is None:
            stats += api.operation_sweep(R.dev, 1, key=key, counter=counter + q, maxnats=maxnats)
        else:
            stats += sweeper(R.dev, key, counter + q, maxnats)
        if learned is not None:
            lst += learned.sweep(counter + q)
        parts = api.energy_cost(R.dev)
        sk, fk = api.capacity_scores(R.dev)
        acc = r * sk / nsweeps if acc is None else acc + r * sk / nsweeps
    return acc, fk, parts, stats, lst


def replica_step(R: ReplicaSet, ctl: ReplicaCtl, it: int, keys: Keys, rng, sys, maxnats: float = 40.0,
                 learned: Learned | None = None, sweeper=None):
    """``replica_step!`` (src/replicas.jl:676-705) of one chain: the weather step of every replica
    (scheduled angle, regime window with a burst, compensated re-dispatch), ``inner`` operation
    sweeps with the mean capacity score, the gates, the drift weights, one Langevin step of the
    shared capacities and the dispatch following on every replica. Returns the step's record."""
    M = R.dev.R
    r = sys.beta_k / sys.beta_v
    th = np.zeros(M)
    wp = np.zeros(M)
    burst = np.zeros(M, dtype=bool)
    ffbs = np.zeros(M, dtype=bool)
    for j in range(M):
        th[j], burst[j] = api.weather_schedule(it, j, M, ctl.theta, ctl.burst_every, ctl.phi)
        # replica_weather! (src/replicas.jl:589-591): a regime move with a burst (or every step in
        # continuous mode) — the window with its probability, the centred FFBS with `ffbs`
        regime = (burst[j] or ctl.burst_every == 0) and (ctl.window > 0 or ctl.ffbs)
        wp[j] = ctl.window_prob if regime and ctl.window > 0 else 0.0
        ffbs[j] = regime and ctl.ffbs
    t0 = time.perf_counter()
    moved = np.zeros(M, dtype=np.int32)
    if (th > 0).any() or (wp > 0).any():
        moved = api.weather_process_step(R.dev, th, wp, window=max(ctl.window, 1), compensate=ctl.comp,
                                         key=keys.weather, counter=it)
    if ffbs.any():
        # weather_process_step's FFBS after the pCN and window moves (src/replicas.jl:195); the
        # anomalies and the weather do not change, so the compensation before it is the same
        api.weather_ffbs(R.dev, mask=ffbs, key=keys.weather, counter=it)
        moved = np.maximum(moved, ffbs.astype(np.int32))
    for j in range(M):
        if burst[j]:
            R.since[j] = 0
            R.window[j] = moved[j] == 2
            R.active[j] = False
            R.gbuf[j].clear()
            R.hist[j].clear()
        else:
            R.since[j] += 1
    Kraw = R.k.copy()
    mask = interior_nodes(Kraw, sys.khi)
    rsk, fk, parts, stats, lst = _sweep_scores(R, ctl.inner, keys.sweep, it * ctl.inner, maxnats, r, learned,
                                               sweeper)
    fkm = masked_means(fk, mask)
    for j in range(M):
        update_gate(R, j, masked_means(rsk[j], mask) / fkm, ctl)
        if R.active[j]:
            R.hist[j].append(rsk[j])
            if len(R.hist[j]) > ctl.probe_win:
                R.hist[j].pop(0)
    wts = drift_weights(R.active)
    xi = rng.standard_normal((NK, sys.N))
    knew, nclip = api.langevin_capacities(R.dev, wts, xi, R.eta_w, ctl.kappa_s, kmin=ctl.kmin, clip=ctl.clip,
                                          box=ctl.box)
    R.k = api.follow_capacities(R.dev, knew, kmin=ctl.kmin)
    tw = time.perf_counter() - t0
    rs = masked_means(np.tensordot(wts, rsk, axes=1), mask)
    return dict(it=it, parts=parts, wts=wts, moved=moved, nclip=nclip, Kraw=Kraw, knew=knew, rs=rs, fk=fkm, tw=tw,
                stats=stats, learned=lst)


def probe_candidate(R: ReplicaSet, ctl: ReplicaCtl):
    """The active replica with a full in-flight window that is longest after its burst (:717-721)."""
    cand = [j for j in range(R.dev.R) if R.active[j] and len(R.hist[j]) >= ctl.probe_win]
    if not cand:
        return None
    return cand[int(np.argmax([R.since[j] for j in cand]))]


def lag_probe(R: ReplicaSet, j: int, probe: api.Replicas, ctl: ReplicaCtl, sys, key, counter: int,
              maxnats: float = 40.0):
    """``lag_probe`` (src/replicas.jl:733-750): sweep a copy of replica j at frozen (k, x) and
    compare its equilibrated score with the replica's in-flight score, per k field."""
    r = sys.beta_k / sys.beta_v
    api.copy_replica(R.dev, j, probe, 0)
    parts_in = api.energy_cost(probe)[0]
    acc, fk, parts = None, None, None
    for q in range(ctl.probe_len):
        api.operation_sweep(probe, 1, key=key, counter=counter + q, maxnats=maxnats)
        if q >= ctl.probe_len - ctl.probe_tail:
            parts = api.energy_cost(probe)[0]
            sk, fk = api.capacity_scores(probe)
            acc = r * sk[0] / ctl.probe_tail if acc is None else acc + r * sk[0] / ctl.probe_tail
    inflight = np.mean(R.hist[j], axis=0)
    mask = interior_nodes(R.k, sys.khi)
    rs_eq, rs_in, fkm = masked_means(acc, mask), masked_means(inflight, mask), masked_means(fk, mask)
    return dict(gap=(rs_in - rs_eq) / fkm, rs_eq=rs_eq, rs_in=rs_in, fk=fkm, op_in=parts_in[1], op_eq=parts[1],
                shed_in=parts_in[4], shed_eq=parts[4])


def adapt_step(R: ReplicaSet, gap, ctl: ReplicaCtl) -> bool:
    """``adapt_step!`` (src/replicas.jl:760-773)."""
    R.gaps.append(np.array(gap, dtype=np.float64))
    if len(R.gaps) < ctl.probe_avg:
        return False
    for f in range(len(R.eta_w)):
        v = [g[f] for g in R.gaps if math.isfinite(g[f])]
        if not v:
            continue
        gm = float(np.mean(v))
        if abs(gm) > ctl.gap_hi:
            R.eta_w[f] = max(ctl.eta_min, R.eta_w[f] * ctl.shrink)
        elif abs(gm) < ctl.gap_lo:
            R.eta_w[f] = min(ctl.eta_max, R.eta_w[f] * ctl.grow)
    R.gaps.clear()
    return True


MONITOR_NAMES = [f"k_{c}" for c in synth.CARRIERS] + ["capex", "opex", "total", "shed_MWh", "penalty", "co2_t"]


def replica_monitor(R: ReplicaSet, parts):
    """Replica means of the monitored quantities (:827): capacity totals per carrier (shared),
    capex, opex, total = capex + opex (src/sampler.jl:34), shed, penalty, CO2."""
    k = R.k
    p = parts.mean(axis=0)
    return np.concatenate([k.sum(axis=1), [p[0], p[1], p[0] + p[1], p[4], p[6], p[7]]])


def init_replicas(sys, k0, M: int, prec: int, ctl: ReplicaCtl, keys: Keys, nsweeps: int, maxnats: float = 40.0):
    """``init_replicas`` (src/replicas.jl:783-795): M replicas at capacities k0, each with an exact
    prior weather draw, the heuristic operation for its weather, and ``nsweeps`` sweeps at fixed k."""
    dev = api.Replicas(sys, M, prec)
    ctl.sweep_options(dev)
    api.sample_weather(dev, key=keys.weather, counter=0xFFFFFFF0)
    _, _, x = dev.download()
    dev.upload(v=synth.heuristic_state(sys, x, k0), k=k0)
    R = ReplicaSet(dev, ctl.eta_w)
    R.k = np.array(k0, dtype=np.float64)
    if nsweeps > 0:
        api.operation_sweep(dev, nsweeps, key=keys.init, counter=0, maxnats=maxnats)
    return R


def chain_start_capacities(sys, kind: str):
    """Capacities of a chain start: ×0.5 (low), ×2 (high) or as set up (src/run.jl:77)."""
    f = {"low": 0.5, "high": 2.0, "heuristic": 1.0}[kind]
    return np.clip(sys.k0 * f, sys.klo, sys.khi)


class BatchedSweeps:
    """The operation sweeps of all chains of a step as one HIP graph (``api.Batch``): every chain's
    thread hands its sweep (key, counter) to ``sweep`` and waits at a barrier; the last one to
    arrive runs the batched sweep for all, while the others wait, so no handle is used by two
    threads at once. A chain that fails aborts the barrier, which ends the other chains' waits."""

    def __init__(self, devs):
        import threading
        self.devs = list(devs)
        self.index = {id(d): c for c, d in enumerate(self.devs)}
        self.batch = api.Batch(self.devs)
        self.req = [None] * len(self.devs)
        self.out = [None] * len(self.devs)
        self.barrier = threading.Barrier(len(self.devs), action=self._run)

    def _run(self):
        keys = [r[0] for r in self.req]
        counters = [r[1] for r in self.req]
        if len({r[2] for r in self.req}) != 1:
            raise ValueError("the chains of a batched sweep use one maxnats")
        st = self.batch.sweeps(1, keys=keys, counters=counters, maxnats=self.req[0][2])
        for c in range(len(self.devs)):
            self.out[c] = st[c]

    def sweep(self, dev, key, counter, maxnats):
        c = self.index[id(dev)]
        self.req[c] = (tuple(key), int(counter), float(maxnats))
        self.barrier.wait()
        return self.out[c]

    def abort(self):
        self.barrier.abort()

    def close(self):
        self.batch.close()


def system_of(cfg: Config):
    """The synthetic system of a configuration: the finest level's nodes and hours per step, the
    coarsest level's cells as countries (when it is coarser than the nodes), K regimes, the cost
    keys and temperatures; with ``co2_price`` = μ the gas unit's marginal cost gains μ·ef (the
    synthetic marginal costs contain no CO2 price)."""
    M0 = cfg.levels[0][0]
    sys = synth.make_system(N=cfg.nodes, L=cfg.lines, T=cfg.steps, dt=cfg.hours, seed=cfg.system_seed, K=cfg.K,
                            ncountry=M0 if 2 <= M0 < cfg.nodes else None)
    mc = sys.mc
    if not math.isnan(cfg.co2_price):
        mc = mc.copy()
        mc[synth.KGAS] = mc[synth.KGAS] + cfg.co2_price * sys.ef
    return dataclasses.replace(sys, c_shed=cfg.c_shed, c_dump=cfg.c_dump, c_viol=cfg.c_viol, beta_v=cfg.beta_v,
                               beta_k=cfg.beta_k, mc=mc)


def run_replicas(cfg: Config, log=None):
    """``run_replicas`` (src/replicas.jl:841-947) on one GPU: ``nchains`` chains of ``replicas``
    weather replicas, ``mlda_warmup`` + ``nsamples`` replica steps, probes and step adaptation
    during the warmup, and the convergence report of the replica-mean quantities over the
    production window, with ESS per second of production wall time."""
    cfg = cfg.check() if isinstance(cfg, Config) else load_config(cfg)
    say = log or (lambda s: None)
    t_setup = time.perf_counter()
    sys = system_of(cfg)
    for note in cfg.notes():
        say(note)
    prec = FP64 if cfg.precision == "fp64" else FP32
    ctl = ReplicaCtl.from_config(cfg)
    nc, M = cfg.nchains, cfg.replicas
    nwarm, nprod = cfg.mlda_warmup, cfg.nsamples
    kinds = [cfg.chain_starts[c % len(cfg.chain_starts)] for c in range(nc)]
    keys = [Keys(cfg.seed, c) for c in range(nc)]
    rngs = [np.random.Generator(np.random.Philox(key=[cfg.seed, 1000 * (c + 1)])) for c in range(nc)]
    sets = []
    for c in range(nc):
        sets.append(init_replicas(sys, chain_start_capacities(sys, kinds[c]), M, prec, ctl, keys[c],
                                  cfg.replica_init_sweeps, cfg.maxnats))
        say(f"chain {c} ({kinds[c]}): {M} replicas initialised ({cfg.replica_init_sweeps} sweeps each)")
    snap = rec_io.read_snapshot(cfg.checkpoint, cfg)
    learned = [None] * nc
    if cfg.learned_proposal:
        for c in range(nc):
            learned[c] = Learned(sets[c].dev, cfg, keys[c])
            if snap is None:
                learned[c].train(cfg.learned_train_steps)
        if snap is None:
            say(f"learned proposal: {cfg.learned_train_steps} training steps per chain")
    # one probe handle per chain (src/replicas.jl:733: a probe slot per chain), so chains share nothing
    probes_h = [api.Replicas(sys, 1, prec) if ctl.probe_every > 0 else None for _ in range(nc)]
    for q in probes_h:
        if q is not None:
            ctl.sweep_options(q)
    traces = [[] for _ in range(nc)]
    records = [[] for _ in range(nc)]
    probes = [[] for _ in range(nc)]
    nprobe = [0] * nc
    it0, restarts, prod_before = 0, 0, 0.0
    if snap is not None:
        # resume (src/replicas.jl:1133-1167): the replicas' states, the driver's gates, buffers,
        # Langevin steps and generator states, the records so far
        it0, restarts = int(snap["it"]), int(snap["restarts"]) + 1
        prod_before = float(snap.get("run/production_s", 0.0))
        for c in range(nc):
            g = f"chains/{c + 1}"
            if str(snap[f"{g}/start"]) != kinds[c] or int(snap[f"{g}/replicas"]) != M:
                raise ValueError(f"snapshot {cfg.checkpoint}: chain {c + 1} does not fit the configuration")
            R = sets[c]
            api.set_weather(R.dev, snap[f"{g}/state/m"], snap[f"{g}/state/z"])
            R.dev.upload(v=snap[f"{g}/state/v"], k=snap[f"{g}/state/k"], x=snap[f"{g}/state/x"])
            R.k = snap[f"{g}/state/kraw"].copy()
            R.eta_w = snap[f"{g}/eta_w"].copy()
            R.since = snap[f"{g}/since"].astype(np.int64)
            R.window = snap[f"{g}/window"].astype(bool)
            R.active = snap[f"{g}/active"].astype(bool)
            R.gbuf, R.hist = rec_io.restore_driver_lists(snap, c)
            R.gaps = [np.array(q) for q in snap.get(f"{g}/driver/gaps", np.zeros((0, 0)))]
            rngs[c].bit_generator.state = rec_io.rng_state_from_json(str(snap[f"{g}/driver/rng"]))
            nprobe[c] = int(snap[f"{g}/driver/nprobe"])
            traces[c], records[c], probes[c] = rec_io.chain_records(snap, c)
            if learned[c] is not None and f"{g}/learned/params" in snap:
                L = learned[c]
                L.net.params(set_to=snap[f"{g}/learned/params"])
                if f"{g}/learned/adam" in snap:
                    L.net.optimizer(set_to=snap[f"{g}/learned/adam"])
                L.ntrain = int(snap[f"{g}/learned/ntrain"])
                L.att, L.acc = (int(q) for q in snap[f"{g}/learned/counts"])
                L.loss = list(snap[f"{g}/learned/loss"])
                L.wall = float(snap[f"{g}/learned/wall"])
        say(f"resumed from {cfg.checkpoint} after step {it0} (restart {restarts})")
    t_init = time.perf_counter()
    batched = BatchedSweeps([R.dev for R in sets]) if cfg.sweep_batch == "graph" else None
    sweeper = batched.sweep if batched is not None else None

    def chain_step(c, it):
        try:
            chain_step_(c, it)
        except BaseException:
            if batched is not None:
                batched.abort()
            raise

    def chain_step_(c, it):
        R = sets[c]
        if (learned[c] is not None and cfg.learned_train_every > 0 and it <= nwarm
                and it % cfg.learned_train_every == 0):
            learned[c].train(cfg.learned_train_more)
        t_step = time.perf_counter()
        st = replica_step(R, ctl, it, keys[c], rngs[c], sys, cfg.maxnats, learned[c], sweeper)
        traces[c].append(replica_monitor(R, st["parts"]))
        rec = rec_io.replica_step_record(R.k, st["parts"], st["wts"], R.active, R.since, st["moved"], st["rs"], st["fk"],
                                         R.eta_w, st["nclip"], time.perf_counter() - t_step, st["tw"], st["learned"])
        # the replicas' weather features at the set-up start's capacities (src/replicas.jl:818, :863)
        rec["features"] = api.replica_features(R.dev, sys.k0).ravel()
        rec.update(it=it, k=R.k.copy(), nclip=int(st["nclip"]))
        records[c].append(rec)
        if it % 25 == 0 or it == nwarm + nprod:
            say(f"chain {c} step {it}/{nwarm + nprod}: total {traces[c][-1][MONITOR_NAMES.index('total')]:.6g}")
        if ctl.probe_every > 0 and it % ctl.probe_every == 0:
            j = probe_candidate(R, ctl)
            if j is not None:
                pr = lag_probe(R, j, probes_h[c], ctl, sys, keys[c].probe, nprobe[c] * ctl.probe_len, cfg.maxnats)
                nprobe[c] += 1
                eta0 = R.eta_w.copy()
                adapted = ctl.adapt and it <= nwarm and adapt_step(R, pr["gap"], ctl)
                probes[c].append(dict(rec_io.probe_record(it, j, adapted, pr, eta0, R.eta_w), j=j, adapted=adapted,
                                      **{k: v for k, v in pr.items() if k not in ("gap", "rs_eq", "rs_in", "fk")}))

    # the chains are independent: one host thread each, so their device work (one HIP stream per
    # handle) overlaps; every chain's results are those of a sequential run (its own keys)
    def snapshot(it, prod_s):
        rec_io.write_snapshot(cfg.checkpoint, cfg, it, restarts, kinds, sets, traces, records, probes, rngs, nprobe,
                              learned, prod_s=prod_s)

    t_prod0 = time.perf_counter() if it0 >= nwarm else None    # resumed inside the production window
    status = 0
    it_last = it0
    t_ckpt = time.perf_counter()
    with ThreadPoolExecutor(max_workers=nc) as ex:
        for it in range(it0 + 1, nwarm + nprod + 1):
            if it == nwarm + 1:
                t_prod0 = time.perf_counter()
            for f in [ex.submit(chain_step, c, it) for c in range(nc)]:
                f.result()
            it_last = it
            if it == nwarm + nprod:
                break
            # step boundary: the stop file (snapshot, remove it, exit code 75) and timed snapshots
            prod_s = prod_before + (time.perf_counter() - t_prod0 if t_prod0 is not None else 0.0)
            if cfg.stop_file and os.path.exists(cfg.stop_file):
                if not cfg.checkpoint:
                    raise ValueError("stop_file needs checkpoint: the snapshot to resume from")
                snapshot(it, prod_s)
                os.remove(cfg.stop_file)
                say(f"stopped after step {it}: snapshot {cfg.checkpoint}")
                status = rec_io.EXIT_CHECKPOINTED
                break
            if cfg.checkpoint and time.perf_counter() - t_ckpt >= 60.0 * cfg.checkpoint_minutes:
                snapshot(it, prod_s)
                t_ckpt = time.perf_counter()
    t_end = time.perf_counter()
    if t_prod0 is None:
        t_prod0 = t_end
    # convergence report on the replica-mean quantities (src/chains.jl:301-310)
    nq = len(MONITOR_NAMES)
    qtr = [[np.array([tr[t][q] for t in range(len(tr))]) for tr in traces] for q in range(nq)]
    good = [all(np.all(np.isfinite(c)) for c in qtr[q]) for q in range(nq)]
    # the steps in the traces (fewer than planned when a stop file ended the run; inside the warmup
    # there is no production window to report on)
    ndone = min(len(tr) for tr in traces)
    w0 = nwarm
    if ndone >= 20:
        wdet = warmup_end([qtr[q] for q in range(nq) if good[q]], threshold=cfg.rhat_screen,
                          every=max(5, ndone // 40))
        w0 = max(nwarm, wdet)
        if w0 >= ndone - 4:
            w0 = nwarm
    summ = {}
    for q in range(nq):
        if good[q] and ndone - w0 >= 4:
            summ[MONITOR_NAMES[q]] = convergence_summary([c[w0:] for c in qtr[q]], rhat_max=cfg.rhat_max,
                                                         ess_min=cfg.ess_min, mcse_max=cfg.mcse_max)
    prod_s = prod_before + (t_end - t_prod0)
    tot = summ.get("total")
    if batched is not None:
        batched.close()
    for R in sets:
        R.dev.close()
    for q in probes_h:
        if q is not None:
            q.close()
    passed = all(q["passed"] for q in summ.values()) if summ else False
    if cfg.out and status == 0:
        rec_io.write_replica_outputs(cfg.out, cfg, sys, ctl, records, probes, restarts, it0 if snap is not None else -1)
        rec_io.write_replica_summary(rec_io.summary_path(cfg.out), ctl, records, probes, kinds, nwarm, passed,
                                     t_end - t_setup, feature_names=api.replica_feature_names(sys))
    lsum = None
    if cfg.learned_proposal:
        lsum = dict(attempted=sum(q.att for q in learned), accepted=sum(q.acc for q in learned),
                    train_steps=sum(len(q.loss) for q in learned),
                    loss_first=[q.loss[0] if q.loss else None for q in learned],
                    loss_last=[q.loss[-1] if q.loss else None for q in learned],
                    train_s=sum(q.wall for q in learned))
    return dict(names=MONITOR_NAMES, traces=traces, records=records, probes=probes, summary=summ, window_start=w0,
                learned=lsum, status=status, restarts=restarts, resumed_at=it0 if snap is not None else -1,
                kinds=kinds, wall=dict(setup=t_init - t_setup, steps=t_end - t_init, production=prod_s),
                steps_per_s=(it_last - it0) / max(t_end - t_init, 1e-12), steps_done=it_last,
                ess_total=None if tot is None else tot["ess_bulk"],
                ess_per_s=None if tot is None or prod_s <= 0 else tot["ess_bulk"] / prod_s,
                # with the learned proposal, the training time counts as production time too
                ess_per_s_with_training=None if tot is None or prod_s <= 0 else
                tot["ess_bulk"] / (prod_s + (lsum["train_s"] if lsum else 0.0)),
                passed=all(s["passed"] for s in summ.values()) if summ else False)


def main(argv=None):
    """``julia run.jl config.toml`` for this driver: run the replica sampler of a TOML
    configuration and exit with 0 (finished) or 75 (stopped by the stop file after a snapshot)."""
    import sys as _sys
    args = _sys.argv[1:] if argv is None else argv
    if len(args) != 1:
        print("usage: python -m esmsampling.replicas config.toml", file=_sys.stderr)
        return 2
    res = run_replicas(load_config(args[0]), log=lambda m: print(m, flush=True))
    print(f"steps {res['steps_done']}, steps/s {res['steps_per_s']:.3f}, ESS/s of total cost {res['ess_per_s']}, "
          f"convergence {'passed' if res['passed'] else 'not passed'}", flush=True)
    return int(res["status"])


if __name__ == "__main__":
    raise SystemExit(main())
