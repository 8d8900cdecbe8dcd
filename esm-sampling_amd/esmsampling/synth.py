"""Seeded synthetic energy system of the benchmark configurations.

The reference reads PyPSA-Eur data from ``data/processed/model.h5`` (schema
``docs/system/data.md:19-35``, loaded by ``src/system.jl:37``), which is not shipped. This module
builds the synthetic system instead:

* network: a random connected graph (a random spanning tree over random positions, each node
  attached to its nearest placed node, plus the shortest random extra edges up to L lines), no
  parallel lines (the reference's angle family assumes none, ``docs/system/device.md:348``);
  reactance log-uniform over [0.01, 0.2] p.u., susceptance b = 1/x normalised to mean 1
  (``docs/system/data.md:28``), thermal limits uniform over [0.5, 3] GW;
* six techs per node: solar, onshore wind, offshore wind (investable renewables as load), gas
  (investable dispatchable), battery (6 h) and H2 store (168 h) (``docs/system/model.md:20-25``);
* weather: 3N logit capacity-factor series, a 2-regime VAR(1) with AR coefficient 0.9, switch
  probability 0.02 per step, innovation std 0.3, regime intercepts and annual harmonics in the
  parameter layout of ``src/weather.jl:461-481`` (B_k = [Φ M_k], Σ_k = σ² I); capacity factor
  = sigmoid(xbar + z) (a declared deviation from the reference's clamp data map,
  ``src/energy.jl:515``; the clamp is kept and is the identity on (0, 1));
* demand: d_n (1 + 0.2 daily sinusoid + N(0, 0.05)) per node and step.

Every number comes from ``numpy.random.default_rng(seed)``, so a seed fixes the system.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

NREN, NSTOR, NK, NV = 3, 2, 6, 4
KGAS, KSTOR = 3, 4
CARRIERS = ("solar", "onwind", "offwind", "gas", "battery", "H2")
VFIELDS = ("theta", "g_gas", "se_battery", "se_H2")


@dataclasses.dataclass
class System:
    N: int
    L: int
    T: int
    dt: float
    line_bus0: np.ndarray      # int32 [L], 0-based
    line_bus1: np.ndarray
    line_b: np.ndarray         # [L]
    line_s: np.ndarray         # [L] MW
    load: np.ndarray           # [N, T] MW
    mc: np.ndarray             # [NK, N] €/MWh
    ef: np.ndarray             # [N] t/MWh (gas)
    capex: np.ndarray          # [NK, N] €/MW/yr
    klo: np.ndarray            # [NK, N]
    khi: np.ndarray            # [NK, N]
    stor_hours: np.ndarray     # [NSTOR, N]
    stor_eta_in: np.ndarray
    stor_eta_out: np.ndarray
    stor_qt: np.ndarray
    c_shed: float = 3000.0
    c_dump: float = 5.0
    c_viol: float = 3000.0
    gamma_k: float = 10.0
    kappa0: float = 1.0
    beta_v: float = 4e-5
    beta_k: float = 4e-8
    # weather model
    K: int = 2
    Ptr: np.ndarray = None     # [K, K]
    pi0: np.ndarray = None     # [K]
    phi: float = 0.9
    sigma: float = 0.3
    sigma0: float = 0.3 / math.sqrt(1 - 0.81)
    mu: np.ndarray = None      # [K, NREN*N, 3]
    xbar: np.ndarray = None    # [NREN*N, 24]
    k0: np.ndarray = None      # [NK, N] heuristic start capacities
    pos: np.ndarray = None     # [N, 2] node positions (graph construction only)
    ncountry: int = 0          # middle partition level of the cell moves (0: none)
    country: np.ndarray = None # [N] int32 in 0..ncountry-1

    def nodelines(self):
        """Lines incident to every node (``LocalModel.nodelines``, src/local.jl:242-245)."""
        out = [[] for _ in range(self.N)]
        for e in range(self.L):
            out[int(self.line_bus0[e])].append(e)
            out[int(self.line_bus1[e])].append(e)
        return out


def _logit(p):
    return np.log(p) - np.log1p(-p)


def _graph(rng, N, L):
    pos = rng.random((N, 2))
    order = rng.permutation(N)
    edges = set()
    placed = [int(order[0])]
    for i in order[1:]:
        i = int(i)
        d = np.sum((pos[placed] - pos[i]) ** 2, axis=1)
        j = placed[int(np.argmin(d))]
        edges.add((min(i, j), max(i, j)))
        placed.append(i)
    if L < N - 1:
        raise ValueError(f"a connected graph on {N} nodes needs at least {N - 1} lines")
    # extra edges: each candidate pairs a random node with one of its 6 nearest neighbours
    d2 = np.sum((pos[:, None, :] - pos[None, :, :]) ** 2, axis=2)
    near = np.argsort(d2, axis=1)[:, 1:7]
    guard = 0
    while len(edges) < L:
        guard += 1
        if guard > 1000 * L:
            raise ValueError("could not place the extra lines")
        a = int(rng.integers(N))
        b = int(near[a, rng.integers(near.shape[1])])
        e = (min(a, b), max(a, b))
        edges.add(e)
    edges = sorted(edges)
    perm = rng.permutation(len(edges))
    edges = [edges[p] for p in perm]
    bus0 = np.array([e[0] for e in edges], dtype=np.int32)
    bus1 = np.array([e[1] for e in edges], dtype=np.int32)
    return pos, bus0, bus1


def _countries(pos, C=None):
    """The middle level of the partition hierarchy (the reference's countries, E.parts,
    src/energy.jl:292): C angular sectors (max(2, N // 5) by default) of the node positions around
    their centroid, of near-equal size, labelled in order of first appearance in node order. Uses
    no random numbers, so it leaves the rest of the system unchanged."""
    N = len(pos)
    C = max(2, N // 5) if C is None else int(C)
    ang = np.arctan2(pos[:, 1] - pos[:, 1].mean(), pos[:, 0] - pos[:, 0].mean())
    order = np.argsort(ang, kind="stable")
    sector = np.empty(N, dtype=np.int64)
    sector[order] = (np.arange(N) * C) // N
    relabel = {}
    for n in range(N):
        relabel.setdefault(int(sector[n]), len(relabel))
    return C, np.array([relabel[int(s)] for s in sector], dtype=np.int32)


def make_system(N: int = 40, L: int = 70, T: int = 8760, dt: float = 1.0, seed: int = 1, K: int = 2,
                ncountry: int | None = None) -> System:
    """The synthetic system with N nodes, L lines and T steps of dt hours (T·dt = 8760 for one year),
    K weather regimes (switch probability 0.02 per step, spread evenly over the other regimes;
    regime intercepts evenly spaced in [−a, a]) and ``ncountry`` countries (max(2, N // 5) by
    default). K and ncountry use no random numbers: at K = 2 the weather is the benchmark
    configurations' 2-regime process."""
    if not 2 <= K <= 4:
        raise ValueError("K: 2 to 4 regimes")
    if ncountry is not None and not 2 <= ncountry <= N:
        raise ValueError("ncountry: 2 to N countries")
    rng = np.random.default_rng(seed)
    pos, bus0, bus1 = _graph(rng, N, L)
    x = np.exp(rng.uniform(math.log(0.01), math.log(0.2), L))
    b = 1.0 / x
    b = b / b.mean()
    s = rng.uniform(500.0, 3000.0, L)

    hod = (np.arange(T) * dt) % 24.0
    dscale = rng.uniform(2000.0, 12000.0, N)
    daily = np.sin(2 * math.pi * (hod - 9.0) / 24.0)
    noise = rng.normal(0.0, 0.05, (N, T))
    load = dscale[:, None] * (1.0 + 0.2 * daily[None, :] + noise)

    mc = np.zeros((NK, N))
    mc[0], mc[1], mc[2] = 0.1, 0.2, 0.3
    mc[3] = rng.uniform(60.0, 90.0, N)
    mc[4], mc[5] = 1.0, 2.0
    ef = np.full(N, 0.37)
    capex = np.zeros((NK, N))
    capex[0] = 40e3
    capex[1] = 90e3
    capex[2] = 160e3
    capex[3] = 50e3
    capex[4] = 70e3
    capex[5] = 110e3
    capex *= rng.uniform(0.9, 1.1, (NK, N))
    peak = load.max(axis=1)
    khi = np.zeros((NK, N))
    khi[0] = rng.uniform(5e3, 30e3, N)
    khi[1] = rng.uniform(5e3, 40e3, N)
    khi[2] = rng.uniform(1e3, 20e3, N)
    khi[3] = 3.0 * peak
    khi[4] = 10e3
    khi[5] = 10e3
    klo = np.zeros((NK, N))
    hours = np.zeros((NSTOR, N))
    hours[0], hours[1] = 6.0, 168.0
    eta_in = np.zeros((NSTOR, N))
    eta_out = np.zeros((NSTOR, N))
    eta_in[0], eta_out[0] = 0.95, 0.95
    eta_in[1], eta_out[1] = 0.7, 0.5
    qt = np.zeros((NSTOR, N))

    # weather: series i = c*N + n
    Ptr = np.full((K, K), 0.02 / (K - 1))
    np.fill_diagonal(Ptr, 0.98)
    pi0 = np.full(K, 1.0 / K)
    phi, sigma = 0.9, 0.3
    sigma0 = sigma / math.sqrt(1.0 - phi * phi)
    nser = NREN * N
    mu = np.zeros((K, nser, 3))
    amp = np.repeat(np.array([0.02, 0.05, 0.05]), N)          # regime intercept contrast per carrier
    seas = np.repeat(np.array([-0.03, 0.03, 0.03]), N)        # annual harmonic (solar peaks in summer)
    for q in range(K):
        mu[q, :, 0] = amp * (2.0 * q / (K - 1) - 1.0)
    mu[:, :, 1] = seas[None, :]
    mu[:, :, 2] = 0.005 * rng.standard_normal(nser)[None, :]
    xbar = np.zeros((nser, 24))
    h = np.arange(24.0)
    sol = np.clip(0.65 * np.sin(math.pi * (h - 6.0) / 12.0), 1e-3, None)
    off = rng.normal(0.0, 0.3, (NREN, N))
    xbar[0:N] = _logit(sol)[None, :] + off[0][:, None]
    xbar[N:2 * N] = _logit(0.3) + off[1][:, None]
    xbar[2 * N:3 * N] = _logit(0.45) + off[2][:, None]

    # heuristic start capacities (state-independent)
    mean_load = load.mean(axis=1)
    k0 = np.zeros((NK, N))
    k0[0] = 0.3 * khi[0]
    k0[1] = 0.3 * khi[1]
    k0[2] = 0.3 * khi[2]
    k0[3] = 1.1 * peak
    k0[4] = 0.1 * mean_load
    k0[5] = 0.05 * mean_load
    return System(N=N, L=L, T=T, dt=float(dt), line_bus0=bus0, line_bus1=bus1, line_b=b, line_s=s,
                  load=load, mc=mc, ef=ef, capex=capex, klo=klo, khi=khi, stor_hours=hours,
                  stor_eta_in=eta_in, stor_eta_out=eta_out, stor_qt=qt, K=K, Ptr=Ptr, pi0=pi0,
                  phi=phi, sigma=sigma, sigma0=sigma0, mu=mu, xbar=xbar, k0=k0, pos=pos,
                  ncountry=_countries(pos, ncountry)[0], country=_countries(pos, ncountry)[1])


def hour_of_day(T: int, dt: float) -> np.ndarray:
    """Hour of day (0..23) of steps 1..T (step 1 starts at hour 0)."""
    return ((np.arange(T) * dt) % 24.0).astype(np.int64)


def weather_from_z(sys: System, z: np.ndarray) -> np.ndarray:
    """Raw node weather X [NREN, N, T, R] = sigmoid(xbar(hour of day) + z) for z [NREN*N, T, R]."""
    hod = hour_of_day(sys.T, sys.dt)
    xb = sys.xbar[:, hod]                              # [nser, T]
    X = 1.0 / (1.0 + np.exp(-(xb[:, :, None] + z)))
    return X.reshape(NREN, sys.N, sys.T, z.shape[-1])


def sample_weather_host(sys: System, R: int, seed: int = 0):
    """Host draw of (m [T, R], z [NREN*N, T, R]) from the weather prior, for synthetic inputs.

    The order of ``sample_weather`` (src/weather.jl:461-481): the regime chain, then the VAR(1)
    recursion with Σ0 at step 1. Uses numpy's generator, not the device's Philox streams (the
    device draws its own with ``esm_weather_init``).
    """
    rng = np.random.default_rng(seed)
    T, nser = sys.T, NREN * sys.N
    m = np.zeros((T, R), dtype=np.int32)
    u = rng.random((T, R))
    m[0] = (u[0] >= sys.pi0[0]).astype(np.int32) if sys.K == 2 else 0
    for t in range(1, T):
        p_stay = sys.Ptr[m[t - 1], m[t - 1]]
        m[t] = np.where(u[t] < p_stay, m[t - 1], 1 - m[t - 1])
    z = np.empty((nser, T, R))
    z[:, 0] = sys.sigma0 * rng.standard_normal((nser, R))
    e = rng.standard_normal((nser, T, R))
    for t in range(1, T):
        ang = 2 * math.pi * (t + 1) / T
        mu = sys.mu[m[t]]                     # [R, nser, 3]
        mean = mu[:, :, 0] + math.cos(ang) * mu[:, :, 1] + math.sin(ang) * mu[:, :, 2]
        z[:, t] = sys.phi * z[:, t - 1] + mean.T + sys.sigma * e[:, t]
    return m, z


def heuristic_state(sys: System, X: np.ndarray, k: np.ndarray | None = None) -> np.ndarray:
    """A start v [NV, N, T, R]: zero angles, gas covering the local net load, storage half full.

    Plays the role of ``heuristic_start`` (src/run.jl:469) for the synthetic system.
    """
    k = sys.k0 if k is None else k
    R = X.shape[-1]
    ren = np.einsum("cn,cntr->ntr", k[:NREN], X)
    net = sys.load[:, :, None] - ren
    v = np.zeros((NV, sys.N, sys.T, R))
    kg = np.maximum(k[KGAS], 1e-9)
    v[1] = np.clip(net / kg[:, None, None], 0.0, 1.0)
    v[2] = 0.5
    v[3] = 0.5
    return v
