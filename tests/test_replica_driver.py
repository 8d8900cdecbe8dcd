"""Host logic of the replica driver (esmsampling.replicas): the schedule, the gate, the step
adaptation and the configuration, against the known answers of test/test_replicas.jl:17-56."""
import math

import numpy as np
import pytest

from esmsampling import api
from esmsampling import replicas as RP


def _ctl(**kw):
    cfg = RP.Config(replicas=4, replica_burst_every=8, replica_burst_angle=0.3, replica_gate_min=2,
                    replica_gate_max=5, replica_gate_win=4, replica_gate_tol=0.1, langevin_eta=1e-3,
                    replica_probe_avg=2)
    ctl = RP.ReplicaCtl.from_config(cfg)
    for k, v in kw.items():
        setattr(ctl, k, v)
    return ctl


def test_burst_schedule_known_answers():
    ctl = _ctl()
    sched = [[it for it in range(1, 17) if api.weather_schedule(it, j, 4, ctl.theta, ctl.burst_every, ctl.phi)[1]]
             for j in range(4)]
    assert sched == [[8, 16], [6, 14], [4, 12], [2, 10]]
    assert api.weather_schedule(8, 0, 4, 0.0, 8, 0.3) == (0.3, True)
    assert api.weather_schedule(7, 0, 4, 0.0, 8, 0.3) == (0.0, False)
    assert api.weather_schedule(8, 0, 4, 0.1, 8, 0.3)[0] == pytest.approx(math.acos(math.cos(0.1) * math.cos(0.3)))
    assert api.weather_schedule(7, 0, 4, 0.1, 8, 0.3)[0] == 0.1


def test_relaxation_indicator_known_answers():
    assert RP.relaxation_indicator([np.array([1.0, 2.0])], 4) == math.inf
    g = [np.array(x) for x in ([1.0, 2.0], [1.0, 2.0], [1.5, 2.0], [1.5, 2.2])]
    assert RP.relaxation_indicator(g, 4) == pytest.approx(0.5)
    # non-finite entries (fields without interior nodes) are skipped
    g = [np.array(x) for x in ([1.0, np.nan], [1.0, np.nan])]
    assert RP.relaxation_indicator(g, 2) == 0.0


def test_gate_sequences_known_answers():
    ctl = _ctl(gate_max=10)
    R = RP.ReplicaSet(None, ctl.eta_w, M=1)
    R.since[0], R.active[0] = 0, False
    acts = []
    for g in ([5.0, 1.0], [2.0, 1.0], [1.0, 1.0], [1.0, 1.0], [1.0, 1.0], [1.0, 1.0], [1.0, 1.0]):
        acts.append(RP.update_gate(R, 0, np.array(g), ctl))
        R.since[0] += 1
    assert acts == [False, False, False, False, False, True, True]
    ctl.gate_max = 5
    R.since[0], R.active[0] = 0, False
    R.gbuf[0].clear()
    acts = []
    for it in range(1, 8):
        acts.append(RP.update_gate(R, 0, np.array([float(it), 0.0]), ctl))
        R.since[0] += 1
    assert acts == [False, False, False, False, False, True, True]
    # continuous mode: always active
    ctl.burst_every = 0
    R.active[0] = False
    assert RP.update_gate(R, 0, np.array([1.0, 1.0]), ctl)


def test_step_adaptation_known_answers():
    ctl = _ctl()
    R = RP.ReplicaSet(None, [1e-3, 1e-3], M=1)
    assert not RP.adapt_step(R, [0.5, 0.01], ctl)
    assert RP.adapt_step(R, [-0.1, 0.03], ctl)            # means 0.2 (shrink) and 0.02 (grow)
    np.testing.assert_allclose(R.eta_w, [0.7e-3, 1.2e-3], rtol=1e-15)
    assert R.gaps == []
    # bounds
    R.eta_w[:] = [2.5e-5, 1.9e-3]
    RP.adapt_step(R, [1.0, 0.0], ctl)
    RP.adapt_step(R, [1.0, 0.0], ctl)
    np.testing.assert_allclose(R.eta_w, [ctl.eta_min, ctl.eta_max])


def test_interior_and_masked_means():
    k = np.array([[10.0, 60.0, 100.0], [0.0, 0.0, 0.0]])
    hi = np.array([[1000.0, 1000.0, 100.5], [10.0, 10.0, 10.0]])
    m = RP.interior_nodes(k, hi)
    assert m.tolist() == [[False, True, False], [False, False, False]]
    mm = RP.masked_means(np.array([[1.0, 2.0, 3.0], [1.0, 1.0, 1.0]]), m)
    assert mm[0] == 2.0 and math.isnan(mm[1])


def test_config_keys():
    c = RP.load_config({"init": "heuristic", "replicas": 16, "nchains": 2, "langevin_kappa": {"solar": 900.0}})
    ctl = RP.ReplicaCtl.from_config(c)
    assert ctl.M == 16 and ctl.kappa_s[0] == 900.0 and ctl.kappa_s[1] == 500.0 and ctl.kappa_s[5] == 100.0
    with pytest.raises(ValueError, match="unknown config key"):
        RP.load_config({"init": "heuristic", "replica_typo": 1})
    assert RP.ReplicaCtl.from_config(RP.load_config({"init": "heuristic", "replica_ffbs": True})).ffbs
    c = RP.load_config({"init": "heuristic", "storage_long": 0.5, "storage_long_moves": 8, "storage_long_hours": 100.0})
    ctl = RP.ReplicaCtl.from_config(c)
    assert (ctl.plong, ctl.long_moves, ctl.long_hours) == (0.5, 8, 100.0)
    d = RP.ReplicaCtl.from_config(RP.load_config({"init": "heuristic"}))
    assert (d.plong, d.long_moves, d.long_hours) == (1.0, 32, 24.0)     # src/run.jl:27-29
    assert (d.xpair, d.xlong) == (0.1, 8)                               # src/run.jl:30-31
    c = RP.load_config({"init": "heuristic", "storage_xpair": 0.0, "storage_xlong": 3})
    assert (RP.ReplicaCtl.from_config(c).xpair, RP.ReplicaCtl.from_config(c).xlong) == (0.0, 3)
    with pytest.raises(ValueError, match="storage_xlong"):
        RP.load_config({"init": "heuristic", "storage_xlong": 300})
    with pytest.raises(ValueError, match="storage_xpair_ball"):
        RP.load_config({"init": "heuristic", "storage_xpair_ball": 2})
    with pytest.raises(ValueError, match="storage_long"):
        RP.load_config({"init": "heuristic", "storage_long_moves": 300})
    c = RP.load_config({"init": "heuristic", "learned_proposal": True, "learned_block_hours": 12, "learned_components": 4})
    assert c.learned_block_hours == 12 and c.learned_components == 4 and c.learned_moves == 1
    assert RP.load_config({"init": "heuristic"}).learned_proposal is False
    for bad in ({"learned_block_hours": 25}, {"learned_components": 9}, {"learned_lr": 0.0}, {"learned_batch": 70000}):
        with pytest.raises(ValueError, match="learned_"):
            RP.load_config({"init": "heuristic", "learned_proposal": True, **bad})


def test_config_toml(tmp_path):
    p = tmp_path / "r.toml"
    p.write_text('sampler = "replicas"\ninit = "heuristic"\nnchains = 1\nreplicas = 4\nmlda_warmup = 3\nnsamples = 5\n'
                 'chain_starts = ["high"]\nlangevin_eta = 1e-4\n')
    c = RP.load_config(str(p))
    assert c.nchains == 1 and c.replicas == 4 and c.chain_starts == ["high"] and c.langevin_eta == 1e-4


# Key sets of the reference's configuration files (configs/*.toml), written out here: the
# replica driver accepts every key of the reference's Config, runs the ones the replica path
# uses, and refuses values it does not build with the reference's message style.
DEVICE40_REPLICA = dict(out="results/device40_replica.h5", basis="haar", storage="soc", renewables="load",
                        measure="physical", levels=[[10, 24], [40, 24]], init="heuristic", use_bias=False,
                        nchains=4, chain_starts=["low", "high"], mlda_warmup=100, nsamples=300,
                        device_backend="gpu", device_sweeps=10, device_sampler="replicas", replicas=4)
CHAINS40 = dict(out="results/chains40.h5", basis="haar", measure="physical", levels=[[10, 24], [40, 24]],
                init="heuristic", use_bias=False, fine_weather=0.05, nchains=4, mlda_warmup=20, nsamples=80)


def test_reference_config_key_sets():
    c = RP.load_config(DEVICE40_REPLICA)     # configs/device40_replica.toml
    assert (c.nodes, c.hours, c.steps, c.lines) == (40, 24.0, 365, 70)
    assert c.measure == "physical" and c.K == 4 and c.levels == [(10, 24), (40, 24)]
    ctl = RP.ReplicaCtl.from_config(c)
    assert ctl.measure == "physical" and ctl.extra == 4.0
    assert "use_bias ignored" not in " ".join(c.notes())
    c = RP.load_config(CHAINS40)             # configs/chains40.toml: fine_weather is not read by the replica path
    assert c.fine_weather == 0.05 and c.sampler == "chains"
    # configs/haar_poc.toml: no init key, so the LP start (not built) -> refused, named
    with pytest.raises(ValueError, match='init = .lp.*not built'):
        RP.load_config(dict(out="results/haar_poc.h5", basis="haar", levels=[[10, 24], [40, 24], [40, 6]],
                            use_bias=False, descent_iters=1500, warmup0=30, mlda_warmup=10, nsamples=30))
    # configs/flow.toml and complete40.toml: the flow basis -> refused, named
    with pytest.raises(ValueError, match="basis = 'flow'"):
        RP.load_config(dict(out="results/flow.h5", basis="flow", levels=[[10, 24], [40, 24], [40, 6]], use_bias=False,
                            descent_iters=0, warmup0=10, mlda_warmup=5, nsamples=20, init="heuristic"))
    for key, val, msg in (("losses", True, "losses"), ("thermal", "soft", "hard"), ("uc", ["gas"], "unit commitment"),
                          ("link_losses", True, "links"), ("brownfield", True, "brownfield"),
                          ("data", "data/processed/model256.h5", "synthetic"), ("storage", "x", "soc"),
                          ("measure", "x", "physical"), ("co2_cap", 1e8, "cap"), ("K", 6, "regimes")):
        with pytest.raises(ValueError, match=msg):
            RP.load_config(dict(DEVICE40_REPLICA, **{key: val}))
    # removed keys: accepted at their one value only (src/run.jl:133-141)
    assert RP.load_config(dict(DEVICE40_REPLICA, kernel="local")).replicas == 4
    with pytest.raises(ValueError, match="no longer supported"):
        RP.load_config(dict(DEVICE40_REPLICA, kernel="flow"))
    # every reference key has a field, with the reference's default
    d = RP.Config()
    assert (d.storage_long, d.storage_long_moves, d.storage_long_hours, d.local_extra, d.K) == (1.0, 32, 24.0, 4.0, 4)
    assert d.init == "lp" and d.sampler == "chains" and d.levels == [(10, 24), (40, 24)] and d.use_bias


def test_levels_set_the_synthetic_system():
    c = RP.load_config(dict(DEVICE40_REPLICA, levels=[[8, 24], [40, 3]]))
    s = RP.system_of(c)
    assert (s.N, s.L, s.T, s.dt) == (40, 70, 2920, 3.0) and s.ncountry == 8 and s.K == 4
    np.testing.assert_allclose(s.Ptr.sum(axis=1), 1.0)
    assert np.all(np.diag(s.Ptr) == 0.98)
    # explicit synthetic keys win; a CO2 price adds μ·ef to the gas marginal cost
    c = RP.load_config(dict(DEVICE40_REPLICA, synthetic_nodes=10, synthetic_lines=16, snapshots=48,
                            resolution_hours=1.0, K=2, co2_price=80.0))
    s = RP.system_of(c)
    base = RP.system_of(RP.load_config(dict(DEVICE40_REPLICA, synthetic_nodes=10, synthetic_lines=16, snapshots=48,
                                            resolution_hours=1.0, K=2)))
    assert (s.N, s.L, s.T, s.K) == (10, 16, 48, 2)
    np.testing.assert_allclose(s.mc[3], base.mc[3] + 80.0 * base.ef)
    assert np.array_equal(s.mc[:3], base.mc[:3])
    with pytest.raises(ValueError, match="levels"):
        RP.load_config(dict(DEVICE40_REPLICA, levels=[[40, 24], [10, 24]]))
