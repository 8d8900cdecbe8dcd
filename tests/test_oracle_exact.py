"""The build's sweep schedule keeps the reference's target (CPU, oracle only).

The oracle's operation sweep restates the schedule of the GPU kernels: atoms in conflict-free
rounds, the random families in a fixed order with the selection shared by all replicas (with the
cross-node storage pairs sxpair at the reference configuration's weight 0.1), the storage
multiscale sweep with probability 1/2 and the long-range storage sweep (with 8 interval storage
pairs sxlong per H2 unit). Its chain is
compared with the reference's own schedule (``Oracle.ref_sweeps``: local_sweep! with the plain
families, per replica and step one shuffled list) on the 3-node network: the means of op, shed,
mean state of charge and a line flow agree, |z| < 4 (test/test_exact.jl:71-106). A kernel that
did not keep the target (a planted bias) fails it.
"""
import numpy as np

import exactness as X
import oracle


def _build_chain(o, s, v, x, burn, nsweeps, key=(22, 1), threads=8):
    k = s.k0
    kw = dict(p_xpair=0.1, xlong_moves=8)   # the storage pairs of the reference configuration
    v, r, F, _ = o.sweeps(v, k, burn, key=key, counter=0, x=x, threads=threads, **kw)
    tr = []
    for q in range(nsweeps):
        v, r, F, _ = o.sweeps(v, k, 1, key=key, counter=burn + q, x=x, r=r, F=F, threads=threads, **kw)
        tr.append(X.observables(o.energy(v, k, x, want_caches=False)[0], v, F))
    return np.array(tr)


def test_build_schedule_keeps_the_reference_target():
    s = X.system()
    R = 64
    v, x = X.start(s, R)
    o = oracle.Oracle(s)
    ref = X.reference_chain(o, s, v, x, burn=300, nsweeps=1000, threads=8)
    test = _build_chain(o, s, v, x, burn=200, nsweeps=300)
    for name, mr, mt, z, er, et in X.zscores(ref, test):
        print(f"{name:10s} ref {mr:.6g} build {mt:.6g} z {z:+.2f} (ESS {er:.0f} / {et:.0f})")
        assert abs(z) < 4.0, (name, mr, mt, z)


def test_exactness_check_detects_a_biased_kernel():
    """The check's power: the same comparison against a chain whose β_v is 10 % higher (a kernel
    that samples another target) fails on op cost."""
    import dataclasses
    s = X.system()
    R = 32
    v, x = X.start(s, R)
    ref = X.reference_chain(oracle.Oracle(s), s, v, x, burn=300, nsweeps=1000, threads=8)
    hot = dataclasses.replace(s, beta_v=1.1 * s.beta_v)
    biased = _build_chain(oracle.Oracle(hot), s, v, x, burn=200, nsweeps=300)
    z = {name: zz for name, _, _, zz, _, _ in X.zscores(ref, biased)}
    assert abs(z["op"]) > 4.0, z
