"""Pins of the FP64 oracle's energy, caches, scores and Langevin step (CPU only).

Each check uses a formulation independent of the oracle's loops, so a dropped term, a wrong
sign or a transposed index fails at least one of them:
* Philox4x32-10 against the Random123 known-answer vectors;
* flows against the incidence-matrix product W Θ, W = diag(b) Dᵀ, and the residual against
  load + D F − Σ A k − k_gas clamp(y) + Σ k (p⁺ − p⁻) written with the incidence matrix D;
* the per-step costs summing to the total operational cost (``cache_cost`` =
  ``energy_cost``, test/test_local.jl:68);
* closed forms of special states (no flows, flat storage);
* the capacity score against central differences of the oracle's own total cost in the node
  capacities (the reference's pin, test/test_replicas.jl:179-203), skipping kinks;
* the Langevin update against the reference's known answers (test/test_replicas.jl:58-84).
"""
import numpy as np
import pytest

import oracle
from conftest import random_state
from esmsampling import synth


def test_philox_known_answers():
    kat = [
        ([0, 0], [0, 0, 0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 2, [0xFFFFFFFF] * 4, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0xA4093822, 0x299F31D0], [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for key, ctr, out in kat:
        assert oracle.philox(key, ctr) == out


def incidence(sys):
    D = np.zeros((sys.N, sys.L))
    for e in range(sys.L):
        D[sys.line_bus0[e], e] = 1.0
        D[sys.line_bus1[e], e] = -1.0
    return D


def soc_terms(sys, s, k, ye, ye1):
    """Storage injection and cost of the SoC form written directly from its docstring
    (src/energy.jl:432-447): d = h (e₁ − e)/Δt − q; charge d/η⁺ if d ≥ 0, else discharge
    min(−d η⁻, 1); running cost Δt mc k (p⁺ + p⁻); box penalty c_viol w h |y − clamp y|;
    beyond the power limit c_viol Δt w (p − 1)₊ (no inflow)."""
    N = sys.N
    h = sys.stor_hours[s][:, None, None]
    ei, eo = sys.stor_eta_in[s][:, None, None], sys.stor_eta_out[s][:, None, None]
    ec, ec1 = np.clip(ye, 0, 1), np.clip(ye1, 0, 1)
    d = h * (ec1 - ec) / sys.dt
    pc = np.where(d >= 0, d / ei, 0.0)
    pm_raw = np.where(d < 0, -d * eo, 0.0)
    pm = np.minimum(pm_raw, 1.0)
    kk = k[:, None, None]
    w = kk + sys.kappa0
    mc = sys.mc[synth.KSTOR + s][:, None, None]
    cost = (sys.dt * mc * kk * (pc + pm) + sys.c_viol * w * h * np.abs(ye - ec)
            + sys.c_viol * sys.dt * w * np.maximum(pc - 1, 0) + sys.c_viol * sys.dt * w * np.maximum(pm_raw - 1, 0))
    return kk * (pc - pm), cost


def test_energy_against_matrix_form(small_system):
    sys = small_system
    R = 3
    v, k, x = random_state(sys, R, seed=11)
    orc = oracle.Oracle(sys)
    parts, ph, r, F = orc.energy(v, k, x)
    kc = np.clip(k, sys.klo, sys.khi)
    D = incidence(sys)
    W = np.diag(sys.line_b) @ D.T                     # L × N
    Fm = np.einsum("ln,ntr->ltr", W, v[0])
    np.testing.assert_allclose(F, Fm, rtol=1e-13, atol=1e-9)
    A = np.clip(x, 0, 1)
    gen_out = np.einsum("cn,cntr->ntr", kc[:3], A) + kc[3][:, None, None] * np.clip(v[1], 0, 1)
    stor_inj = np.zeros_like(gen_out)
    stor_cost = np.zeros_like(gen_out)
    for s in range(2):
        ye = v[2 + s]
        inj, c = soc_terms(sys, s, kc[4 + s], ye, np.roll(ye, -1, axis=1))
        stor_inj += inj
        stor_cost += c
    rm = np.maximum(sys.load, 0)[:, :, None] + np.einsum("nl,ltr->ntr", D, Fm) - gen_out + stor_inj
    np.testing.assert_allclose(r, rm, rtol=1e-12, atol=1e-7)
    dt = sys.dt
    line_pen = sys.c_viol * dt * np.maximum(np.abs(Fm) - sys.line_s[:, None, None], 0).sum(axis=0)
    gen_cost = dt * (np.einsum("cn,cntr->tr", sys.mc[:3] * kc[:3], A)
                     + np.einsum("n,ntr->tr", sys.mc[3] * kc[3], np.clip(v[1], 0, 1)))
    gas_pen = sys.c_viol * dt * np.einsum("n,ntr->tr", kc[3] + sys.kappa0, np.abs(v[1] - np.clip(v[1], 0, 1)))
    bal = dt * (sys.c_shed * np.maximum(rm, 0) + sys.c_dump * np.maximum(-rm, 0)).sum(axis=0)
    ph_m = line_pen + gen_cost + gas_pen + stor_cost.sum(axis=0) + bal
    np.testing.assert_allclose(ph, ph_m, rtol=1e-11)
    np.testing.assert_allclose(parts[:, 1], ph.sum(axis=0), rtol=1e-12)
    # parts: shed/dump in MWh, capex with the box penalty γ_k capex |k − kc|
    np.testing.assert_allclose(parts[:, 4], dt * np.maximum(rm, 0).sum(axis=(0, 1)), rtol=1e-12)
    np.testing.assert_allclose(parts[:, 5], dt * np.maximum(-rm, 0).sum(axis=(0, 1)), rtol=1e-12)
    capex = (sys.capex * k).sum() + sys.gamma_k * (sys.capex * np.abs(k - kc)).sum()
    np.testing.assert_allclose(parts[:, 0], capex, rtol=1e-13)
    np.testing.assert_allclose(parts[:, 7], dt * np.einsum("n,ntr->r", sys.ef * kc[3], np.clip(v[1], 0, 1)), rtol=1e-12)
    # op = gen + storage + c_shed shed + c_dump dump + penalty
    np.testing.assert_allclose(parts[:, 1], parts[:, 2] + parts[:, 3] + sys.c_shed * parts[:, 4]
                               + sys.c_dump * parts[:, 5] + parts[:, 6], rtol=1e-12)


def test_energy_closed_form_no_flows(small_system):
    """θ = 0 (no flows), flat states of charge (no storage power), utilizations inside the box:
    op = Σ Δt (mc A k + mc_g k_g y) + Σ Δt (c_shed r₊ + c_dump r₋) with r = load − A k − k_g y."""
    sys = small_system
    R = 2
    rng = np.random.default_rng(5)
    v = np.zeros((4, sys.N, sys.T, R))
    v[1] = rng.uniform(0, 1, (sys.N, sys.T, R))
    v[2] = 0.3
    v[3] = 0.7
    x = rng.uniform(0, 1, (3, sys.N, sys.T, R))
    k = sys.k0.copy()
    parts, ph, r, F = oracle.Oracle(sys).energy(v, k, x)
    assert np.all(F == 0)
    out = np.einsum("cn,cntr->ntr", k[:3], x) + k[3][:, None, None] * v[1]
    rr = sys.load[:, :, None] - out
    np.testing.assert_allclose(r, rr, rtol=1e-12, atol=1e-8)
    op = (np.einsum("cn,cntr->r", sys.mc[:3] * k[:3], x) + np.einsum("n,ntr->r", sys.mc[3] * k[3], v[1])
          + sys.c_shed * np.maximum(rr, 0).sum(axis=(0, 1)) + sys.c_dump * np.maximum(-rr, 0).sum(axis=(0, 1)))
    np.testing.assert_allclose(parts[:, 1], op * sys.dt, rtol=1e-12)
    assert np.all(parts[:, 6] == 0)


def test_storage_power_limit_and_efficiency(small_system):
    """One node with a rising state of charge: charge d/η_in above 1 is priced at c_viol Δt w (p − 1)."""
    sys = small_system
    R = 1
    v = np.zeros((4, sys.N, sys.T, R))
    v[1] = 0.5
    v[2] = 0.5
    v[3] = 0.5
    h = sys.stor_hours[0, 0]
    v[2, 0, 1, 0] = 0.5 + 1.5 * sys.dt / h        # d = 1.5 at step 1 (0-based 0 → 1 charges), then discharge
    x = np.full((3, sys.N, sys.T, R), 0.3)
    k = sys.k0.copy()
    o = oracle.Oracle(sys)
    p1, ph1, r1, _ = o.energy(v, k, x)
    v0 = v.copy()
    v0[2, 0, 1, 0] = 0.5
    p0, ph0, r0, _ = o.energy(v0, k, x)
    kb = k[4, 0]
    ei, eo = sys.stor_eta_in[0, 0], sys.stor_eta_out[0, 0]
    pc = 1.5 / ei
    # step 0: charging at pc: injection +k pc at node 0; step 1: discharge min(1.5 η_out, 1)
    assert r1[0, 0, 0] - r0[0, 0, 0] == pytest.approx(kb * pc, rel=1e-12)
    assert r1[0, 1, 0] - r0[0, 1, 0] == pytest.approx(-kb * min(1.5 * eo, 1.0), rel=1e-12)
    w = kb + sys.kappa0
    pen = sys.c_viol * sys.dt * w * (pc - 1) + sys.c_viol * sys.dt * w * max(1.5 * eo - 1, 0)
    assert p1[0, 6] - p0[0, 6] == pytest.approx(pen, rel=1e-12)


def test_scores_against_central_differences(small_system):
    """ŝ_k = −β_v ∂C/∂k against central differences of the oracle's op in each node capacity."""
    sys = small_system
    R = 2
    v, k, x = random_state(sys, R, seed=21)
    o = oracle.Oracle(sys)
    sk, fk, sa = o.scores(v, k, x)
    checked = 0
    for f in range(6):
        for n in range(sys.N):
            if not (sys.klo[f, n] < k[f, n] < sys.khi[f, n]):
                assert np.all(sk[:, f, n] == 0)
                continue
            hstep = 1e-4 * max(1.0, k[f, n])
            kp, km = k.copy(), k.copy()
            kp[f, n] += hstep
            km[f, n] -= hstep
            op_p = o.energy(v, kp, x, want_caches=False)[0][:, 1]
            op_m = o.energy(v, km, x, want_caches=False)[0][:, 1]
            fd = -sys.beta_v * (op_p - op_m) / (2 * hstep)
            # skip kinks: the one-sided differences must agree
            op0 = o.energy(v, k, x, want_caches=False)[0][:, 1]
            right = (op_p - op0) / hstep
            left = (op0 - op_m) / hstep
            smooth = np.abs(right - left) <= 1e-6 * (np.abs(right) + np.abs(left) + 1e-9)
            for j in range(R):
                if smooth[j]:
                    assert sk[j, f, n] == pytest.approx(fd[j], rel=1e-5, abs=1e-9 * sys.beta_v * sa[j, f, n])
                    checked += 1
    assert checked >= 10 * R
    # the capex force: β_k capex inside the box, β_k capex (1 + γ_k) above it
    kc = np.clip(k, sys.klo, sys.khi)
    np.testing.assert_allclose(fk, sys.beta_k * sys.capex * (1 + sys.gamma_k * np.sign(k - kc)), rtol=1e-14)


def test_physical_measure_scores_against_logtarget_differences(small_system):
    """With measure = "physical": −fk + ŝ_k is the gradient of the log target at fixed (v, x),
    logtarget = −β_k K − β_v op + log J (src/hierarchy.jl:90; the reference's pin
    test/test_replicas.jl:179-203 with its replica config's measure), against central
    differences of the oracle's capex, op and log J in every node capacity, kinks skipped.
    log J itself: zero for the utilization measure, invariant under moving k beyond its box
    (it uses the clamped capacity), and additive over units: lowering one gas capacity from k to
    k' changes it by exactly T log((k' + κ0)/(k + κ0))."""
    sys = small_system
    R = 2
    v, k, x = random_state(sys, R, seed=23)
    o = oracle.Oracle(sys)
    assert o.log_measure(k, "utilization") == 0.0
    k2 = k.copy()
    k2[synth.KGAS, 0] = 0.5 * k[synth.KGAS, 0] + 0.5 * sys.klo[synth.KGAS, 0]
    assert o.log_measure(k2) - o.log_measure(k) == pytest.approx(
        sys.T * np.log((k2[synth.KGAS, 0] + sys.kappa0) / (k[synth.KGAS, 0] + sys.kappa0)), rel=1e-12)
    k3 = k.copy()
    k3[:, 1] = sys.khi[:, 1] + 5.0
    k4 = k.copy()
    k4[:, 1] = sys.khi[:, 1]
    assert o.log_measure(k3) == pytest.approx(o.log_measure(k4), rel=1e-15)
    sk, fk, sa = o.scores(v, k, x, measure="physical")
    sk_u, _, _ = o.scores(v, k, x)
    assert np.all(sk[:, :synth.KGAS] == sk_u[:, :synth.KGAS])      # renewables: the utilization measure

    def logtarget(kk):
        p = o.energy(v, kk, x, want_caches=False)[0]
        return -sys.beta_k * p[:, 0] - sys.beta_v * p[:, 1] + o.log_measure(kk)

    checked = 0
    L0 = logtarget(k)
    for f in range(6):
        for n in range(sys.N):
            if not (sys.klo[f, n] < k[f, n] < sys.khi[f, n]):
                continue
            hstep = 1e-4 * max(1.0, k[f, n])
            kp, km = k.copy(), k.copy()
            kp[f, n] += hstep
            km[f, n] -= hstep
            Lp, Lm = logtarget(kp), logtarget(km)
            fd = (Lp - Lm) / (2 * hstep)
            right, left = (Lp - L0) / hstep, (L0 - Lm) / hstep
            smooth = np.abs(right - left) <= 1e-6 * (np.abs(right) + np.abs(left) + 1e-9)
            for j in range(R):
                if smooth[j]:
                    scale = sa[j, f, n] + abs(fk[f, n]) + 1e-12
                    assert -fk[f, n] + sk[j, f, n] == pytest.approx(fd[j], rel=1e-5, abs=1e-9 * scale)
                    checked += 1
    assert checked >= 10 * R


def _langevin_one_field(sys, k, g, xi, eta=1e-3, ks=100.0, kmin=1.0, clip=np.inf, box=False, hi=None):
    """The oracle's Langevin step on field 0 only, with the drift g set through fk = −g, sk = 0."""
    import copy
    s2 = copy.copy(sys)
    N = len(k)
    assert N <= sys.N
    kk = np.full((6, sys.N), 10.0)
    kk[0, :N] = k
    khi = sys.khi.copy()
    if hi is not None:
        khi[0, :N] = hi
    s2.khi = khi
    o = oracle.Oracle(s2)
    fk = np.zeros((6, sys.N))
    fk[0, :N] = -np.asarray(g)
    sk = np.zeros((1, 6, sys.N))
    xi_full = np.zeros((6, sys.N))
    xi_full[0, :N] = xi
    etav = np.zeros(6)
    etav[0] = eta
    knew, nc = o.langevin(kk, sk, fk, np.ones(1), xi_full, etav, np.full(6, ks), kmin=kmin, clip=clip, box=box)
    assert np.all(knew[1:] == kk[1:])
    return knew[0, :N], nc


def test_langevin_known_answers(small_system):
    """test/test_replicas.jl:58-84."""
    sys = small_system
    k = np.array([0.0, 50.0, 1000.0, 5000.0])
    g = np.array([1e-3, -2e-3, 5e-4, 0.0])
    ks, eta = 100.0, 1e-3
    xi = np.array([0.1, -0.2, 0.3, -0.4])
    k1, nc = _langevin_one_field(sys, k, g, xi, eta=eta, ks=ks)
    w = np.log(k + ks)
    w1 = w + eta * ((k + ks) * g + 1) + np.sqrt(2 * eta) * xi
    wmin = np.log(1.0 + ks)
    w1 = np.where(w1 < wmin, 2 * wmin - w1, w1)
    np.testing.assert_allclose(k1, np.exp(w1) - ks, rtol=1e-14)
    assert nc == 0 and np.all(k1 >= 1.0 - 1e-12)
    kr, _ = _langevin_one_field(sys, [2.0], [0.0], [-50.0], eta=eta, ks=ks)
    xx = np.log(2.0 + ks) + eta - np.sqrt(2 * eta) * 50
    assert np.log(kr[0] + ks) == pytest.approx(2 * wmin - xx, rel=1e-13)
    kc, nc = _langevin_one_field(sys, [1000.0, 1000.0], [1.0, -1.0], [0.0, 0.0], eta=eta, ks=ks, clip=0.02)
    assert nc == 2
    np.testing.assert_allclose(np.log(kc + ks), np.log(1000.0 + ks) + np.array([0.02, -0.02]), rtol=1e-13)
    kb, _ = _langevin_one_field(sys, [1090.0, 1200.0], [0.0, 0.0], [5.0, 5.0], eta=eta, ks=ks, box=True,
                                hi=[1100.0, 1100.0])
    wmax = np.log(1100.0 + ks)
    x1 = np.log(1090.0 + ks) + eta + np.sqrt(2 * eta) * 5
    assert np.log(kb[0] + ks) == pytest.approx(2 * wmax - x1, rel=1e-13) and kb[0] < 1100
    assert np.log(kb[1] + ks) == pytest.approx(np.log(1200.0 + ks) + eta + np.sqrt(2 * eta) * 5, rel=1e-13)


def test_weights_known_answers():
    """test/test_replicas.jl:49-55 (host-side weights of the drift)."""
    from esmsampling.api import tilt_weights, uniform_weights
    phis = [0.7, -1.3, 0.2, 2.1]
    act = [True, False, True, True]
    assert list(tilt_weights(0.0, phis, act)) == list(uniform_weights(act)) == [1 / 3, 0.0, 1 / 3, 1 / 3]
    w = tilt_weights(2.0, phis, act)
    assert w.sum() == pytest.approx(1) and w[1] == 0 and w[3] / w[0] == pytest.approx(np.exp(2 * (2.1 - 0.7)))
    assert list(uniform_weights([False] * 4)) == [0.25] * 4
