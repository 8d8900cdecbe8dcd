"""The learned dispatch-block proposal on the GPU (esm_mlp_*) against its CPU restatement and the
exact target.

* the MFMA forward pass against the FP64 network with the device's BF16 rounding points;
* the loss and the gradient of the backward pass (MFMA GEMMs, split-K atomics) against the CPU
  backward pass;
* one learned move of every (node, block, replica), FP64 state: the MH log-ratio from the
  oracle's gas-atom ΔC and the mixture density of the device's outputs and proposals, and the
  accept decisions from the same Philox uniforms;
* exactness: with one-hour blocks the target factorizes over (node, step); after learned moves
  alone, each (node, step)'s values over 2048 replicas follow exp(−β ΔC(y)) (quadrature of the
  oracle's gas atom, KS), so the MH correction leaves the target unchanged;
* what the trained network gains: maximum-likelihood training on the sampler's states lowers the
  loss and raises the acceptance rate of the learned moves over the untrained network.
"""
import numpy as np
import pytest

import mlp_oracle as mo
import oracle
from esmsampling import FP32, FP64, api, synth

pytestmark = pytest.mark.gpu


def _perturbed(net, seed, scale=0.05):
    rng = np.random.default_rng(seed)
    p = net.params()
    p = p + rng.normal(0, scale, p.shape) * np.abs(p).mean()
    net.params(set_to=p)
    return net.layers(net.params())


@pytest.mark.parametrize("ncomp", [4, 8])   # 320 outputs (256-wide tiles), 576 (192-wide tiles)
def test_forward_matches_oracle(ncomp):
    print(f"mlp forward ncomp={ncomp}", flush=True)
    s = synth.make_system(N=6, L=8, T=48, dt=1.0, seed=3)
    dev = api.Replicas(s, 2, FP64)
    net = api.LearnedProposal(dev, block_hours=24, ncomp=ncomp, key=(3, 5))
    layers = _perturbed(net, 1, scale=2.0)
    X = np.random.default_rng(2).normal(0, 1, (700, net.din))
    out = net.forward(X)
    ref, _ = mo.forward(layers, X)
    scale = np.abs(ref).max()
    assert scale > 0
    # BF16 activations: an FP32 sum that rounds to the other BF16 neighbour moves a few outputs
    assert np.max(np.abs(out - ref)) <= 2e-2 * scale
    assert np.median(np.abs(out - ref)) <= 2e-3 * scale
    dev.close()


@pytest.mark.parametrize("ncomp", [2, 4, 8])   # 144, 320 and 576 outputs: 192-column passes, masked or full
def test_fused_forward_is_the_layered_forward_bit_for_bit(ncomp):
    """The fused forward (one kernel per 128-row tile, activations in LDS) against the one-kernel-
    per-layer form: identical outputs for row counts that leave a partial tile."""
    print(f"mlp fused forward ncomp={ncomp}", flush=True)
    s = synth.make_system(N=6, L=8, T=48, dt=1.0, seed=3)
    dev = api.Replicas(s, 2, FP64)
    net = api.LearnedProposal(dev, block_hours=24, ncomp=ncomp, key=(3, 5))
    _perturbed(net, 4, scale=2.0)
    for M in (1, 700, 4096 + 77):
        X = np.random.default_rng(M).normal(0, 1, (M, net.din))
        net.options(fused=True)
        a = net.forward(X)
        net.options(fused=False)
        b = net.forward(X)
        assert np.array_equal(a, b), (M, np.max(np.abs(a - b)))
        assert np.all(np.isfinite(a)) and np.abs(a).max() > 0
    net.options(fused=True)
    with pytest.raises(Exception, match="fused is 0 or 1"):
        api.check(dev.lib.esm_mlp_options(dev.handle, 2), "esm_mlp_options")
    dev.close()


def test_gradient_matches_oracle():
    print("mlp gradient", flush=True)
    s = synth.make_system(N=6, L=8, T=48, dt=1.0, seed=3)
    dev = api.Replicas(s, 2, FP64)
    net = api.LearnedProposal(dev, block_hours=24, ncomp=4, key=(4, 6))
    layers = _perturbed(net, 3, scale=2.0)
    rng = np.random.default_rng(4)
    M = 1500                                            # several split-K chunks of the weight GEMMs
    X = rng.normal(0, 1, (M, net.din))
    mu0 = rng.uniform(0, 1, (M, net.B))
    y = mu0 + rng.normal(0, 0.2, (M, net.B))
    loss, g = net.gradient(X, y, mu0)
    out, acts = mo.forward(layers, X)
    loss_ref, dg = mo.head_grad(out, mo.f32(mu0), mo.f32(y), net.K, net.B, net.dout)
    gr = mo.flatten(mo.backward(layers, acts, dg))
    assert loss == pytest.approx(loss_ref, rel=1e-2)
    cos = float(g @ gr / (np.linalg.norm(g) * np.linalg.norm(gr)))
    assert cos > 0.999, cos
    assert np.linalg.norm(g - gr) <= 5e-2 * np.linalg.norm(gr)
    dev.close()


@pytest.mark.parametrize("N,L,B", [(6, 8, 2), (6, 8, 24), (40, 70, 24)])
def test_learned_move_ratio_and_decisions_fp64(N, L, B):
    """The learned move's FP64 MH log-ratio against −β ΔC of the oracle's gas atoms plus the
    oracle's proposal densities, and identical decisions from the same Philox uniforms; on a
    6-node network and on the 40-node / 70-line network of the learned-proposal workload."""
    print(f"mlp move fp64 N={N} B={B}", flush=True)
    s = synth.make_system(N=N, L=L, T=48, dt=1.0, seed=5)
    R = 3
    rng = np.random.default_rng(6)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    v[1] = np.clip(v[1] + rng.normal(0, 0.05, v[1].shape), -0.02, 1.02)
    dev = api.Replicas(s, R, FP64)
    dev.upload(v=v, k=s.k0, x=x)
    net = api.LearnedProposal(dev, block_hours=B, ncomp=3, key=(5, 7))
    layers = _perturbed(net, 5)
    v0, k0, x0 = dev.download()
    r0, F0 = dev.caches()
    key, ctr = (5, 6), 3
    st, out, yp, lr = net.propose(key=key, counter=ctr, trace=True)
    v1, _, _ = dev.download()
    K, nb = net.K, net.nb
    X, Y, MU0 = mo.features(s, v0, k0, x0, r0, F0, B, R)
    ref, _ = mo.forward(layers, X)
    assert np.max(np.abs(out - ref)) <= 2e-2 * np.abs(ref).max()
    lq_old = mo.logq(out, mo.f32(MU0), Y, K, B)[0]      # the state's own y
    lq_new = mo.logq(out, mo.f32(MU0), yp, K, B)[0]
    o = oracle.Oracle(s)
    decided = 0
    for i in range(net.n_inst):
        j = i % R
        n, d = divmod(i // R, nb)
        dc = sum(o.move_cost(v0, k0, r0, F0, j, 0, 0, n, d * B + b, [yp[i, b] - v0[1, n, d * B + b, j]])[0]
                 for b in range(B))
        lr_ref = -s.beta_v * dc + lq_old[i] - lq_new[i]
        assert abs(lr[i] - lr_ref) <= 1e-4 * (1.0 + abs(lr_ref)), (i, lr[i], lr_ref)
        w = oracle.philox(key, (j, n * nb + d, 0x5B000000, ctr))
        ua = ((int(w[2]) << 32 | int(w[3])) >> 11) * 2.0 ** -53
        acc = np.array_equal(v1[1, n, d * B:d * B + B, j], yp[i])
        if abs(np.log(ua) - lr_ref) > 1e-3:
            assert acc == (ua < np.exp(min(0.0, lr_ref))), (i, ua, lr_ref)
            decided += 1
        if not acc:
            assert np.array_equal(v1[1, n, d * B:d * B + B, j], v0[1, n, d * B:d * B + B, j])
    assert decided >= net.n_inst - 2 and st[0] == net.n_inst and st[1] < st[0]
    if B == 2:
        assert st[1] > 0
    # the caches follow the accepted blocks
    rd, _ = dev.caches()
    _, _, r_, _ = o.energy(v1, k0, x0)
    np.testing.assert_allclose(rd, r_, atol=1e-9 * np.abs(s.load).max())
    dev.close()


def test_learned_move_keeps_the_target():
    """One-hour blocks: each (node, step) is one instance and the target factorizes; after learned
    moves alone the replicas' values at a (node, step) follow exp(−β ΔC(y)) of the oracle's gas atom."""
    print("mlp exactness", flush=True)
    s = synth.make_system(N=4, L=4, T=6, dt=1.0, seed=7)
    R = 2048
    rng = np.random.default_rng(8)
    x1 = rng.uniform(0.1, 0.9, (3, s.N, s.T, 1))
    v1 = synth.heuristic_state(s, x1)
    v = np.repeat(v1, R, axis=-1)
    x = np.repeat(x1, R, axis=-1)
    dev = api.Replicas(s, R, FP32)
    dev.upload(v=v, k=s.k0, x=x)
    net = api.LearnedProposal(dev, block_hours=1, ncomp=2, key=(8, 9))
    acc = att = 0
    for c in range(60):
        st = net.propose(key=(13, 14), counter=c)
        acc += st[1]
        att += st[0]
    assert acc > 0.08 * att, acc / att
    vd, _, _ = dev.download()
    o = oracle.Oracle(s)
    _, _, r1, F1 = o.energy(v1, s.k0, x1)
    for (n, tau) in [(0, 0), (1, 3), (2, 5), (3, 2)]:
        y0 = v1[1, n, tau, 0]
        grid = np.linspace(-0.3, 1.3, 40001)
        lp = -s.beta_v * o.move_cost(v1, s.k0, r1, F1, 0, 0, 0, n, tau, grid - y0)
        w = np.exp(lp - lp.max())
        cdf = np.cumsum(w) / w.sum()
        ys = np.sort(vd[1, n, tau, :])
        F_emp = np.arange(1, R + 1) / R
        F_q = np.interp(ys, grid, cdf)
        ks = np.max(np.abs(F_emp - F_q))
        assert ks < 0.05, (n, tau, ks)
    dev.close()


def test_training_lowers_loss_and_raises_acceptance():
    """What the trained network gains over the untrained one: maximum likelihood on the states of
    the plain sampler, then the acceptance rate of one learned move of every 24-hour block on fresh
    states (ten plain sweeps later), the same states for both networks. A mixture of K = 8
    Gaussians per hour fitted exactly to this target's hourly shape (a shedding cliff below the
    requirement, an exponential tail above) accepts ≈ 0.4 of the 24-hour blocks; the untrained
    network (sd 0.2 around the requirement) practically none."""
    print("mlp training", flush=True)
    s = synth.make_system(N=10, L=16, T=240, dt=1.0, seed=9)
    R = 256
    rng = np.random.default_rng(10)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    dev = api.Replicas(s, R, FP32)
    dev.upload(v=v, k=s.k0, x=x)
    api.operation_sweep(dev, 20, key=(2, 2), counter=0)      # states of the plain sampler
    net = api.LearnedProposal(dev, block_hours=24, ncomp=8, key=(10, 11))
    loss = net.train(1000, batch=4096, lr=2e-4, key=(12, 12))
    assert np.all(np.isfinite(loss)) and loss[-50:].mean() < loss[:50].mean() - 10.0, (loss[:5], loss[-5:])
    api.operation_sweep(dev, 10, key=(2, 3), counter=0)      # fresh states
    vs, ks, xs = dev.download()
    twin = api.Replicas(s, R, FP32)
    twin.upload(v=vs, k=ks, x=xs)
    base = api.LearnedProposal(twin, block_hours=24, ncomp=8, key=(10, 11))   # the same initial network
    st_t = net.propose(key=(3, 3), counter=0)
    st_u = base.propose(key=(3, 3), counter=0)
    rate_t, rate_u = st_t[1] / st_t[0], st_u[1] / st_u[0]
    print(f"acceptance: trained {rate_t:.3f}, untrained {rate_u:.3f}; loss {loss[:50].mean():.2f} -> "
          f"{loss[-50:].mean():.2f}", flush=True)
    assert st_t[0] == s.N * (s.T // 24) * R
    assert rate_u < 0.01 and rate_t > 0.08
    dev.close()
    twin.close()


def test_training_is_bitwise_reproducible():
    """Two networks trained from the same start on the same states with the same key end with the
    same parameters and losses, bit for bit: the gradient reductions (K-split partials, bias
    column sums, loss) are summed in a fixed order, whatever order the workgroups finish in."""
    print("mlp training reproducible", flush=True)
    s = synth.make_system(N=10, L=16, T=96, dt=1.0, seed=9)
    R = 64
    rng = np.random.default_rng(3)
    x = rng.uniform(0.05, 0.95, (3, s.N, s.T, R))
    v = synth.heuristic_state(s, x)
    out = []
    for _ in range(2):
        dev = api.Replicas(s, R, FP32)
        dev.upload(v=v, k=s.k0, x=x)
        net = api.LearnedProposal(dev, block_hours=24, ncomp=4, key=(4, 4))
        loss = net.train(40, batch=8192, lr=1e-3, key=(5, 5))
        out.append((net.params(), loss))
        dev.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
