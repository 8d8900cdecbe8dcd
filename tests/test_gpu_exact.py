"""Chain-marginal exactness of the composed GPU kernel (test/test_exact.jl:71-106).

On the seeded 3-node network at fixed capacities and weather, the GPU's operation sweep as the
replica sampler runs it — atoms in conflict-free rounds, the random families in a fixed order
with the move selection shared by the replicas and the cross-node storage pairs sxpair (weight
0.1), the storage multiscale sweep with probability 1/2, the long-range storage sweep with 8
interval storage pairs sxlong per H2 unit, and (FP32 run) one learned move of every (node, step,
replica) after every sweep, its network trained on the chain's own states — must
reproduce the equilibrium of the reference's own schedule (the oracle's ``ref_sweeps``: atoms
plus the plain families, per replica and step one shuffled list, FP64): the means of op cost,
shed, mean state of charge and a line flow agree, |z| < 4.
"""
import numpy as np
import pytest

import exactness as X
import oracle
from esmsampling import FP32, FP64, api

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(600)
@pytest.mark.parametrize("prec,learned", [(FP64, False), (FP32, True)])
def test_gpu_kernel_keeps_the_reference_target(prec, learned):
    print(f"exactness prec={prec} learned={learned}: reference chain (oracle)", flush=True)
    s = X.system()
    R = 64
    v, x = X.start(s, R)
    o = oracle.Oracle(s)
    ref = X.reference_chain(o, s, v, x, burn=300, nsweeps=1000)
    print("  GPU chain", flush=True)
    dev = api.Replicas(s, R, prec)
    dev.upload(v=v, k=s.k0, x=x)
    api.storage_cross_options(dev, p_xpair=0.1, nxlong=8)   # the reference configuration's storage pairs
    key = (23, 1 + int(learned))
    api.operation_sweep(dev, 200, key=key, counter=0)
    net = None
    if learned:
        # one-step blocks: at daily steps the shedding cliff of a block is narrower than the
        # proposal's smallest standard deviation, so longer blocks are rarely accepted
        net = api.LearnedProposal(dev, block_hours=1, ncomp=4, key=(23, 7))
        net.train(500, batch=4096, lr=1e-3, key=(23, 8))
    tr, att, acc = [], 0, 0
    for q in range(1000):
        api.operation_sweep(dev, 1, key=key, counter=200 + q)
        if net is not None:
            st = net.propose(key=(23, 9), counter=q)
            att += int(st[0])
            acc += int(st[1])
        parts = api.energy_cost(dev)
        vd, _, _ = dev.download()
        _, Fd = dev.caches()
        tr.append(X.observables(parts, vd, Fd))
    test = np.array(tr)
    dev.close()
    if net is not None:
        print(f"  learned moves accepted {acc} of {att}", flush=True)
        assert acc > 0
    for name, mr, mt, z, er, et in X.zscores(ref, test):
        print(f"  {name:10s} reference {mr:.6g} GPU {mt:.6g} z {z:+.2f} (ESS {er:.0f} / {et:.0f})", flush=True)
        assert abs(z) < 4.0, (name, mr, mt, z)
