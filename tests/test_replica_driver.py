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
    c = RP.load_config({"replicas": 16, "nchains": 2, "langevin_kappa": {"solar": 900.0}})
    ctl = RP.ReplicaCtl.from_config(c)
    assert ctl.M == 16 and ctl.kappa_s[0] == 900.0 and ctl.kappa_s[1] == 500.0 and ctl.kappa_s[5] == 100.0
    with pytest.raises(ValueError, match="unknown config key"):
        RP.load_config({"replica_typo": 1})
    with pytest.raises(ValueError, match="FFBS"):
        RP.load_config({"replica_ffbs": True})
    c = RP.load_config({"storage_long": 0.5, "storage_long_moves": 8, "storage_long_hours": 100.0})
    ctl = RP.ReplicaCtl.from_config(c)
    assert (ctl.plong, ctl.long_moves, ctl.long_hours) == (0.5, 8, 100.0)
    d = RP.ReplicaCtl.from_config(RP.load_config({}))
    assert (d.plong, d.long_moves, d.long_hours) == (1.0, 32, 24.0)     # src/run.jl:27-29
    with pytest.raises(ValueError, match="sxpair"):
        RP.load_config({"storage_xpair": 0.1})
    with pytest.raises(ValueError, match="sxpair"):
        RP.load_config({"storage_xlong": 8})
    with pytest.raises(ValueError, match="storage_long"):
        RP.load_config({"storage_long_moves": 300})
    c = RP.load_config({"learned_proposal": True, "learned_block_hours": 12, "learned_components": 4})
    assert c.learned_block_hours == 12 and c.learned_components == 4 and c.learned_moves == 1
    assert RP.load_config({}).learned_proposal is False
    for bad in ({"learned_block_hours": 25}, {"learned_components": 9}, {"learned_lr": 0.0}, {"learned_batch": 70000}):
        with pytest.raises(ValueError, match="learned_"):
            RP.load_config({"learned_proposal": True, **bad})


def test_config_toml(tmp_path):
    p = tmp_path / "r.toml"
    p.write_text('sampler = "replicas"\nnchains = 1\nreplicas = 4\nmlda_warmup = 3\nnsamples = 5\n'
                 'chain_starts = ["high"]\nlangevin_eta = 1e-4\n')
    c = RP.load_config(str(p))
    assert c.nchains == 1 and c.replicas == 4 and c.chain_starts == ["high"] and c.langevin_eta == 1e-4
