"""Convergence diagnostics and effective sample sizes of the sampler's traces (host, FP64).

Restates src/convergence.jl:1-147 (rank-normalized split-R̂, bulk and tail ESS, MCSE, warmup
detection; Vehtari, Gelman, Simpson, Carpenter, Bürkner 2021) and the single-chain `ess` of
src/diagnostics.jl:8-28. These are measurement code: they read traces that the replica driver
recorded and never touch the device.

The inverse normal CDF of the rank normalization (the reference calls SpecialFunctions'
`erfcinv`, src/convergence.jl:36) is Wichura's algorithm AS 241 (PPND16, Applied Statistics 37,
1988, 477-484), accurate to about 1e-16 relative.
"""
from __future__ import annotations

import math

import numpy as np

# AS 241 PPND16 coefficients (Wichura 1988), numerator a/c/e and denominator b/d/f, low order first.
_A = (3.3871328727963666080e0, 1.3314166789178437745e2, 1.9715909503065514427e3, 1.3731693765509461125e4,
      4.5921953931549871457e4, 6.7265770927008700853e4, 3.3430575583588128105e4, 2.5090809287301226727e3)
_B = (1.0, 4.2313330701600911252e1, 6.8718700749205790830e2, 5.3941960214247511077e3,
      2.1213794301586595867e4, 3.9307895800092710610e4, 2.8729085735721942674e4, 5.2264952788528545610e3)
_C = (1.42343711074968357734e0, 4.63033784615654529590e0, 5.76949722146069140550e0, 3.64784832476320460504e0,
      1.27045825245236838258e0, 2.41780725177450611770e-1, 2.27238449892691845833e-2, 7.74545014278341407640e-4)
_D = (1.0, 2.05319162663775882187e0, 1.67638483018380384940e0, 6.89767334985100004550e-1,
      1.48103976427480074590e-1, 1.51986665636164571966e-2, 5.47593808499534494600e-4, 1.05075007164441684324e-9)
_E = (6.65790464350110377720e0, 5.46378491116411436990e0, 1.78482653991729133580e0, 2.96560571828504891230e-1,
      2.65321895265761230930e-2, 1.24266094738807843860e-3, 2.71155556874348757815e-5, 2.01033439929228813265e-7)
_F = (1.0, 5.99832206555887937690e-1, 1.36929880922735805310e-1, 1.48753612908506148525e-2,
      7.86869131145613259100e-4, 1.84631831751005468180e-5, 1.42151175831644588870e-7, 2.04426310338993978564e-15)


def _poly(c, x):
    y = np.zeros_like(x)
    for a in reversed(c):
        y = y * x + a
    return y


def norm_ppf(p):
    """Φ⁻¹(p) for p in (0, 1) (AS 241); ±inf at 0 and 1."""
    p = np.asarray(p, dtype=np.float64)
    q = p - 0.5
    out = np.empty_like(p)
    central = np.abs(q) <= 0.425
    if central.any():
        qc = q[central]
        r = 0.180625 - qc * qc
        out[central] = qc * _poly(_A, r) / _poly(_B, r)
    tail = ~central
    if tail.any():
        qt = q[tail]
        pm = np.minimum(p[tail], 1.0 - p[tail])
        with np.errstate(divide="ignore", invalid="ignore"):     # p = 0 or 1: r = inf, replaced below
            r = np.sqrt(-np.log(pm))
            x = np.empty_like(r)
            near = r <= 5.0
            rn = r[near] - 1.6
            x[near] = _poly(_C, rn) / _poly(_D, rn)
            rf = r[~near] - 5.0
            x[~near] = _poly(_E, rf) / _poly(_F, rf)
        x = np.where(pm <= 0.0, np.inf, x)
        out[tail] = np.where(qt < 0, -x, x)
    return out


def _chains(chains):
    return [np.asarray(c, dtype=np.float64).ravel() for c in chains]


def split_chains(chains):
    """Every chain split into halves; the second half of an odd chain drops its first draw
    (src/convergence.jl:6-16)."""
    out = []
    for c in _chains(chains):
        h = len(c) // 2
        out.append(c[:h])
        out.append(c[len(c) - h:])
    return out


def _average_ranks(x):
    """1-based ranks with ties given their average rank (src/convergence.jl:21-33)."""
    S = len(x)
    p = np.argsort(x, kind="stable")
    xs = x[p]
    r = np.empty(S)
    starts = np.flatnonzero(np.r_[True, xs[1:] != xs[:-1]])
    ends = np.r_[starts[1:], S]
    for i, j in zip(starts, ends):
        r[p[i:j]] = (i + 1 + j) / 2.0
    return r


def rank_normalize(chains):
    """z = Φ⁻¹((rank − 3/8)/(S + 1/4)) over the pooled draws (src/convergence.jl:19-44)."""
    cs = _chains(chains)
    x = np.concatenate(cs) if cs else np.zeros(0)
    S = len(x)
    z = norm_ppf((_average_ranks(x) - 0.375) / (S + 0.25))
    out, o = [], 0
    for c in cs:
        out.append(z[o:o + len(c)])
        o += len(c)
    return out


def rhat_basic(chains):
    """Between/within-variance R̂ of equal-length chains (src/convergence.jl:47-59)."""
    cs = _chains(chains)
    M = len(cs)
    n = min(len(c) for c in cs) if cs else 0
    if M < 2 or n < 2:
        return math.nan
    X = np.stack([c[:n] for c in cs])
    ms = X.mean(axis=1)
    vs = X.var(axis=1, ddof=1)
    W = vs.mean()
    B = n * ms.var(ddof=1)
    if W == 0:
        return 1.0 if B == 0 else math.inf
    vplus = (n - 1) / n * W + B / n
    return math.sqrt(vplus / W)


def rhat_rank(chains):
    """Rank-normalized split-R̂: the larger of the bulk and the folded version (:62-68)."""
    sc = split_chains(chains)
    bulk = rhat_basic(rank_normalize(sc))
    med = np.median(np.concatenate(sc))
    tail = rhat_basic(rank_normalize([np.abs(c - med) for c in sc]))
    return max(bulk, tail)


def autocov(x):
    """Autocovariances at lags 0..n−1, biased (denominator n), by direct sums (:71-76)."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    y = x - x.mean()
    return np.correlate(y, y, mode="full")[n - 1:] / n


def _geyer_tau(rho, n):
    """τ = −1 + 2 Σ_k P_k over the initial positive, monotone sequence of pair sums
    P_k = ρ_{2k} + ρ_{2k+1}."""
    tau, prev, k = -1.0, math.inf, 0
    while 2 * k + 1 < n:
        pk = rho[2 * k] + rho[2 * k + 1]
        if pk <= 0:
            break
        pk = min(pk, prev)
        tau += 2 * pk
        prev = pk
        k += 1
    return tau


def ess_multi(chains):
    """Multi-chain ESS (Stan's estimator, src/convergence.jl:82-108): autocorrelations combined
    through var⁺, Geyer's initial monotone sequence; chains cut to the shortest."""
    cs = _chains(chains)
    M = len(cs)
    n = min(len(c) for c in cs)
    if n < 4:
        return float(M * n)
    cs = [c[:n] for c in cs]
    acs = np.stack([autocov(c) for c in cs])
    W = float(np.mean([c.var(ddof=1) for c in cs]))
    if W == 0:
        return float(M * n)
    ms = np.array([c.mean() for c in cs])
    vplus = (n - 1) / n * W + (ms.var(ddof=1) if M > 1 else 0.0)
    rho = 1.0 - (W - acs.mean(axis=0)) / vplus
    rho[0] = 1.0
    tau = max(_geyer_tau(rho, n), 1.0 / math.log10(M * n))
    return M * n / tau


def ess_bulk(chains):
    """Bulk ESS: ess_multi of the rank-normalized split chains (:111)."""
    return ess_multi(rank_normalize(split_chains(chains)))


def _quantile_type7(x, q):
    return float(np.quantile(x, q))


def ess_tail(chains):
    """Tail ESS: the smaller ESS of the indicators of the 5 % and 95 % quantiles (:114-123)."""
    sc = split_chains(chains)
    x = np.concatenate(sc)
    vals = []
    for q in (0.05, 0.95):
        qv = _quantile_type7(x, q)
        vals.append(ess_multi([(c <= qv).astype(np.float64) for c in sc]))
    return min(vals)


def mcse_rel(chains):
    """MCSE of the mean relative to the posterior sd, 1/√ESS (:126-129)."""
    e = ess_multi(split_chains(chains))
    return 1.0 / math.sqrt(max(e, 1.0))


def convergence_summary(chains, rhat_max=1.01, ess_min=400.0, mcse_max=0.05):
    """Diagnostics of one quantity over several chains and whether it passes (:136-147)."""
    cs = _chains(chains)
    x = np.concatenate(cs)
    if np.all(x == x[0]):
        return dict(rhat=1.0, ess_bulk=float(len(x)), ess_tail=float(len(x)), mcse_rel=0.0, mean=float(x[0]),
                    sd=0.0, passed=True, constant=True)
    r = rhat_rank(cs)
    eb, et = ess_bulk(cs), ess_tail(cs)
    mr = mcse_rel(cs)
    ok = math.isfinite(r) and r < rhat_max and eb >= ess_min and et >= ess_min and mr <= mcse_max
    return dict(rhat=r, ess_bulk=eb, ess_tail=et, mcse_rel=mr, mean=float(x.mean()), sd=float(x.std(ddof=1)),
                passed=bool(ok), constant=False)


def warmup_end(traces, threshold=1.05, every=10, minlen=20):
    """First sweep from which the rank-normalized split-R̂ of every quantity over the second half
    of the sweeps so far stays below ``threshold`` (src/convergence.jl:150-171).
    ``traces[q][k]`` is the trace of quantity q in chain k. Returns 0 if never."""
    n = min(len(c) for tr in traces for c in tr)
    ok = []
    for m in range(minlen, n + 1, every):
        r = 1.0
        for tr in traces:
            head = np.concatenate([np.asarray(c[:m], dtype=np.float64) for c in tr])
            if np.all(head == np.asarray(tr[0], dtype=np.float64)[0]):
                continue
            r = max(r, rhat_rank([np.asarray(c[m // 2:m], dtype=np.float64) for c in tr]))
        ok.append(m if r < threshold else 0)
    fails = [i for i, v in enumerate(ok) if v == 0]
    if not fails:
        return 0 if not ok else ok[0] // 2
    last = fails[-1]
    if last == len(ok) - 1:
        return 0
    return ok[last + 1] // 2


def ess(x):
    """Single-chain ESS with Geyer's initial monotone sequence (src/diagnostics.jl:8-28)."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    if n < 4:
        return float(n)
    mu = x.mean()
    v = np.sum((x - mu) ** 2) / n
    if v == 0:
        return float(n)
    rho = autocov(x) / v
    s, prev, k = 0.0, math.inf, 0
    while 2 * k + 1 < n:
        g = rho[2 * k] + rho[2 * k + 1]
        if g <= 0:
            break
        g = min(g, prev)
        s += g
        prev = g
        k += 1
    tau = -1 + 2 * s
    return n / max(tau, 1.0 / n)
