"""Edge cases of the per-sweep path on the GPU against the oracle (FP64): one replica, the
shortest horizons (two steps: each parity class one step; three steps: an odd cycle whose last
step's storage terms wrap around to the first), a two-node network with one line, gas units
without capacity (the pair families skip units below κmin; the atoms still sample them under the
box penalty alone), the caches after the sweep; the learned proposal on a
horizon that its blocks do not divide (the remaining hours are never proposed) and with blocks
of the whole horizon."""
import numpy as np
import pytest

import oracle
from esmsampling import FP64, api, synth

pytestmark = pytest.mark.gpu


def _sweep_vs_oracle(N, L, T, R, seed=3, zero_gas=()):
    s = synth.make_system(N=N, L=L, T=T, dt=1.0, seed=seed)
    rng = np.random.default_rng(seed + 1)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    v[0] = rng.normal(0, 200.0, (s.N, s.T, R))
    k = s.k0.copy()
    for n in zero_gas:
        k[3, n] = 0.0
    dev = api.Replicas(s, R, FP64)
    dev.upload(v=v, k=k, x=x)
    st = api.operation_sweep(dev, 2, key=(3, 4), counter=1)
    vd, _, _ = dev.download()
    rd, Fd = dev.caches()
    v_, r_, F_, st_ = oracle.Oracle(s).sweeps(v, k, 2, key=(3, 4), counter=1, x=x, threads=2)
    dev.close()
    assert np.array_equal(st[:, :4], st_[:, :4])
    np.testing.assert_allclose(vd[1:], v_[1:], atol=1e-9)
    np.testing.assert_allclose(vd[0], v_[0], atol=1e-9 * max(np.abs(v_[0]).max(), 1.0))
    np.testing.assert_allclose(rd, r_, atol=1e-9 * np.abs(s.load).max())
    np.testing.assert_allclose(Fd, F_, atol=1e-9 * max(np.abs(F_).max(), 1.0))
    return s, v, vd, st


@pytest.mark.parametrize("N,L,T,R", [(6, 8, 4, 1), (5, 6, 2, 3), (5, 6, 3, 3), (2, 1, 6, 2)])
def test_sweep_edge_shapes_match_oracle(N, L, T, R):
    print(f"edge sweep N={N} L={L} T={T} R={R}", flush=True)
    _sweep_vs_oracle(N, L, T, R)


def test_sweep_without_gas_capacity_matches_oracle():
    print("edge sweep zero gas capacity", flush=True)
    s, v, vd, st = _sweep_vs_oracle(6, 8, 6, 2, seed=5, zero_gas=(0, 3))
    # a unit without capacity still has its utilization sampled (under the box penalty alone,
    # w = κ + κ0), but produces nothing: the residual at its node does not depend on it
    assert not np.array_equal(vd[1, [0, 3]], v[1, [0, 3]])
    k = s.k0.copy()
    k[3, [0, 3]] = 0.0
    x = np.random.default_rng(6).uniform(0.05, 0.95, (3, s.N, s.T, 2))
    o = oracle.Oracle(s)
    _, _, r_a, _ = o.energy(vd, k, x)
    vb = vd.copy()
    vb[1, [0, 3]] = 0.5
    _, _, r_b, _ = o.energy(vb, k, x)
    assert np.array_equal(r_a[[0, 3]], r_b[[0, 3]])


def test_learned_proposal_block_edges():
    """Blocks that do not divide the horizon leave its last hours alone; a block of the whole
    horizon is one instance per node and replica."""
    print("edge learned blocks", flush=True)
    s = synth.make_system(N=4, L=5, T=10, dt=1.0, seed=6)
    R = 3
    rng = np.random.default_rng(7)
    x = rng.uniform(0.1, 0.9, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    dev = api.Replicas(s, R, FP64)
    dev.upload(v=v, k=s.k0, x=x)
    net = api.LearnedProposal(dev, block_hours=4, ncomp=2, key=(1, 2))
    assert net.nb == 2 and net.n_inst == s.N * 2 * R
    net.train(300, batch=24, lr=3e-3, key=(2, 2))
    acc = 0
    for c in range(40):
        acc += int(net.propose(key=(4, 4), counter=c)[1])
    vd, _, _ = dev.download()
    assert acc > 0
    assert np.array_equal(vd[1, :, 8:, :], v[1, :, 8:, :])     # hours 8, 9: no block
    assert not np.array_equal(vd[1, :, :8, :], v[1, :, :8, :])
    others = vd.copy()
    others[1] = v[1]
    assert np.array_equal(others, v)                         # only the gas utilization moves
    net = api.LearnedProposal(dev, block_hours=10, ncomp=1, key=(1, 3))
    assert net.nb == 1 and net.n_inst == s.N * R
    st = net.propose(key=(5, 5), counter=0)
    assert st[0] == s.N * R
    with pytest.raises(Exception, match="block_hours"):
        api.LearnedProposal(dev, block_hours=11)
    dev.close()
