"""The chain-marginal exactness check of the reference (test/test_exact.jl:39-106), restated on a
seeded synthetic 3-node network: a chain with some kernel and a reference chain (atoms plus the
plain families in the reference's own shuffled order, ``Oracle.ref_sweeps``) at fixed
capacities and weather must agree on the means of op cost, shed, mean state of charge and one
line's mean flow within Monte Carlo error, |z| < 4 with an ESS-based standard error
(``compare``, :55-64). Both chains run many independent replicas; the trace of an observable is
its replica mean per sweep, whose ESS (src/diagnostics.jl:8-28) covers the correlation between
sweeps and, for kernels whose replicas share their move selection, between replicas.
"""
import numpy as np

from esmsampling import synth
from esmsampling.diagnostics import ess

NAMES = ("op", "shed", "mean SoC", "line flow")


def system():
    """Three nodes on a triangle at daily resolution, 12 steps (a small horizon, as the
    reference's 3-node sub-network at 5-day steps)."""
    return synth.make_system(N=3, L=3, T=12, dt=24.0, seed=2)


def start(s, R, seed=0):
    """Fixed weather (one draw, the same in every replica) and the heuristic start."""
    rng = np.random.default_rng(seed)
    x1 = rng.uniform(0.1, 0.9, (3, s.N, s.T, 1))
    x = np.repeat(x1, R, axis=-1)
    return synth.heuristic_state(s, x), x


def observables(parts, v, F):
    """[4, R]: op, shed (MWh), the mean state of charge of both storage carriers, the mean flow
    of line 0 (b (θ_a − θ_b), test/test_exact.jl:46-53)."""
    return np.stack([parts[:, 1], parts[:, 4], v[2:].mean(axis=(0, 1, 2)), F[0].mean(axis=0)])


def zscores(ref, test):
    """Per observable z of the difference of the means of two traces [S, 4, R], as ``compare``
    (test/test_exact.jl:55-64). The reference chain's replicas are independent chains: its
    standard error is the spread of their time averages over √R, which holds however slowly an
    observable mixes (atoms alone move the states of charge slowly). The tested kernel's
    replicas share their move selection: its standard error is ESS-based on the replica-mean
    trace (src/diagnostics.jl:8-28)."""
    out = []
    R = ref.shape[2]
    for i in range(ref.shape[1]):
        r, t = ref[:, i, :], test[:, i, :].mean(axis=1)
        se_r = r.mean(axis=0).std(ddof=1) / np.sqrt(R)
        se = np.sqrt(se_r ** 2 + t.var(ddof=1) / max(ess(t), 2.0))
        z = (t.mean() - r.mean()) / max(se, 1e-12 * abs(r.mean()) + 1e-12)
        out.append((NAMES[i], r.mean(), t.mean(), float(z), r.var(ddof=1) / max(se_r ** 2, 1e-300), ess(t)))
    return out


def reference_chain(o, s, v, x, burn, nsweeps, threads=16, key=(21, 1)):
    """The reference chain's trace [nsweeps, 4, R] after ``burn`` sweeps."""
    k = s.k0
    v, r, F, _ = o.ref_sweeps(v, k, burn, key=key, counter=0, x=x, threads=threads)
    tr = []
    for q in range(nsweeps):
        v, r, F, _ = o.ref_sweeps(v, k, 1, key=key, counter=burn + q, x=x, r=r, F=F, threads=threads)
        parts = o.energy(v, k, x, want_caches=False)[0]
        tr.append(observables(parts, v, F))
    return np.array(tr)
