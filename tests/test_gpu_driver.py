"""The replica driver on the device: the replica copy of the lag probe (esm_copy_replica) and a
short run of run_replicas with bursts, gates, probes and the convergence report."""
import numpy as np
import pytest

from conftest import random_state
from esmsampling import FP64, api, synth
from esmsampling import replicas as RP

pytestmark = pytest.mark.gpu


def test_copy_replica_matches_source():
    print("copy replica fp64", flush=True)
    s = synth.make_system(N=10, L=16, T=48, dt=1.0, seed=8)
    R = 3
    v, k, x = random_state(s, R, seed=21, spread=False)
    src = api.Replicas(s, R, FP64)
    api.sample_weather(src, key=(4, 4), counter=1)
    src.upload(v=v, k=k)
    dst = api.Replicas(s, 1, FP64)
    api.copy_replica(src, 1, dst, 0)
    vs, ks, xs = src.download()
    vd, kd, xd = dst.download()
    np.testing.assert_array_equal(vd[..., 0], vs[..., 1])
    np.testing.assert_array_equal(xd[..., 0], xs[..., 1])
    np.testing.assert_array_equal(kd, ks)
    ms, zs = api.weather_state(src)
    md, zd = api.weather_state(dst)
    np.testing.assert_array_equal(md[:, 0], ms[:, 1])
    np.testing.assert_array_equal(zd[..., 0], zs[..., 1])
    ps = api.energy_cost(src)
    pd = api.energy_cost(dst)
    np.testing.assert_allclose(pd[0], ps[1], rtol=1e-12)
    # the copy sweeps on its own: the source is unchanged
    api.operation_sweep(dst, 1, key=(1, 1), counter=0)
    np.testing.assert_array_equal(src.download()[0], vs)
    with pytest.raises(Exception, match="out of range"):
        api.copy_replica(src, 3, dst, 0)
    src.close()
    dst.close()


def test_run_replicas_short():
    print("run_replicas 2 chains x 4 replicas", flush=True)
    cfg = RP.load_config(dict(init="heuristic", nchains=2, replicas=4, mlda_warmup=10, nsamples=14, replica_init_sweeps=2,
                              replica_burst_every=4, replica_gate_min=1, replica_gate_max=2, replica_gate_min_window=1,
                              replica_gate_max_window=3, replica_gate_win=2, replica_probe_every=3,
                              replica_probe_len=4, replica_probe_tail=2, replica_probe_win=1, replica_probe_avg=1,
                              replica_window=8, synthetic_nodes=10, synthetic_lines=16, snapshots=48,
                              resolution_hours=1.0, precision="fp64", seed=3))
    res = RP.run_replicas(cfg)
    assert len(res["traces"]) == 2 and len(res["traces"][0]) == 24
    tr = np.array(res["traces"])
    assert np.all(np.isfinite(tr))
    # the capacities are shared within a chain and move by the Langevin step
    k0 = res["records"][0][0]["k"]
    assert not np.array_equal(k0, res["records"][0][-1]["k"])
    # capex + opex = total
    i = RP.MONITOR_NAMES.index
    np.testing.assert_allclose(tr[..., i("total")], tr[..., i("capex")] + tr[..., i("opex")], rtol=1e-12)
    assert "total" in res["summary"] and res["ess_total"] > 0 and res["ess_per_s"] > 0
    assert sum(len(p) for p in res["probes"]) > 0
    for p in res["probes"][0]:
        assert np.all(np.isfinite(p["rs_eq"][np.isfinite(p["fk"])]))
    # the replicas' weather features, stacked replica by replica: finite, regime fractions sum to 1
    K = cfg.K
    for rec in res["records"][1]:
        F = rec["features"].reshape(4, 3 + 5 + K)
        assert np.all(np.isfinite(F)) and np.allclose(F[:, 7:7 + K].sum(axis=1), 1.0) and np.all(F[:, 1:4] <= 1.0)
    # bursts reset the gates: some steps have inactive replicas
    assert any(not r["active"].all() for r in res["records"][0])
    for rec in res["records"][0]:
        w = rec["weights"]
        assert w.sum() == pytest.approx(1.0) and np.all(w[~rec["active"]] == 0) or not rec["active"].any()


def test_run_replicas_learned_proposal():
    """The driver with the learned proposal: trained before the first step and during the warmup
    (never in the production window), one learned sweep after every operation sweep, its counts
    in the records, and the report with the training time counted."""
    print("run_replicas learned 2 chains x 8 replicas", flush=True)
    cfg = RP.load_config(dict(init="heuristic", nchains=2, replicas=8, mlda_warmup=10, nsamples=10, replica_init_sweeps=4,
                              replica_probe_every=0, synthetic_nodes=10, synthetic_lines=16, snapshots=96,
                              resolution_hours=1.0, precision="fp32", seed=5, learned_proposal=True,
                              learned_train_steps=300, learned_train_every=5, learned_train_more=50,
                              learned_components=4))
    res = RP.run_replicas(cfg)
    ls = res["learned"]
    per_step = 10 * (96 // 24) * 8
    assert ls["attempted"] == 2 * 20 * per_step
    assert 0 < ls["accepted"] < ls["attempted"]
    assert ls["train_steps"] == 2 * (300 + 2 * 50)          # warmup steps 5 and 10
    assert all(b < a for a, b in zip(ls["loss_first"], ls["loss_last"]))
    for rec in res["records"][0]:
        assert rec["learned"][0] == per_step
    assert np.all(np.isfinite(np.array(res["traces"])))
    assert res["ess_per_s_with_training"] < res["ess_per_s"]


def test_chains_in_threads_match_a_chain_alone():
    """The chains run in host threads, each on its own handle and stream: chain 0 of a three-chain
    run is bitwise the chain of a one-chain run with the same seed (its own keys, no shared state)."""
    print("run_replicas threads vs alone", flush=True)
    base = dict(init="heuristic", resolution_hours=1.0, replicas=4, mlda_warmup=6, nsamples=6, replica_init_sweeps=2, replica_burst_every=4,
                replica_probe_every=3, replica_probe_len=3, replica_probe_tail=1, replica_probe_win=1,
                replica_probe_avg=1, replica_gate_min=1, replica_gate_max=2, replica_gate_min_window=1,
                replica_gate_max_window=3, replica_gate_win=2, synthetic_nodes=10, synthetic_lines=16, snapshots=48,
                precision="fp64", seed=11, learned_proposal=True, learned_train_steps=20, learned_components=2)
    a = RP.run_replicas(RP.load_config(dict(base, nchains=3, sweep_batch="threads")))
    b = RP.run_replicas(RP.load_config(dict(base, nchains=1, sweep_batch="threads")))
    assert np.array_equal(np.array(a["traces"][0]), np.array(b["traces"][0]))
    for ra, rb in zip(a["records"][0], b["records"][0]):
        assert np.array_equal(ra["k"], rb["k"]) and np.array_equal(ra["learned"], rb["learned"])
    # the chains' sweeps as one captured graph per step: the same run, bitwise
    print("run_replicas graph-batched sweeps vs threads", flush=True)
    g = RP.run_replicas(RP.load_config(dict(base, nchains=3, sweep_batch="graph")))
    for c in range(3):
        assert np.array_equal(np.array(a["traces"][c]), np.array(g["traces"][c]))
        for ra, rg in zip(a["records"][c], g["records"][c]):
            assert np.array_equal(ra["k"], rg["k"]) and np.array_equal(ra["learned"], rg["learned"])


def test_stop_file_snapshot_and_resume(tmp_path):
    """A run stopped by the stop file in the production window (snapshot, exit code 75) and
    resumed from its snapshot continues as the run without a stop: the same traces, capacities
    and Langevin steps (FP64), the records of both parts, the outputs of the finished run."""
    print("stop file, snapshot, resume", flush=True)
    base = dict(init="heuristic", resolution_hours=1.0, nchains=2, replicas=4, mlda_warmup=10, nsamples=20, replica_init_sweeps=2, replica_burst_every=4,
                replica_probe_every=6, replica_probe_len=3, replica_probe_tail=1, replica_probe_win=1,
                replica_probe_avg=1, replica_gate_min=1, replica_gate_max=2, replica_gate_min_window=1,
                replica_gate_max_window=3, replica_gate_win=2, synthetic_nodes=10, synthetic_lines=16, snapshots=48,
                precision="fp64", seed=13)
    a = RP.run_replicas(RP.load_config(dict(base, out=str(tmp_path / "a" / "out.npz"))))
    ck, stop = tmp_path / "ck.npz", tmp_path / "STOP"

    def log(msg):
        if "step 25/" in msg:
            stop.write_text("")

    cfg = RP.load_config(dict(base, checkpoint=str(ck), stop_file=str(stop), out=str(tmp_path / "b" / "out.npz")))
    b1 = RP.run_replicas(cfg, log=log)
    assert b1["status"] == 75 and b1["steps_done"] == 25 and ck.exists() and not stop.exists()
    b2 = RP.run_replicas(cfg)
    assert b2["status"] == 0 and b2["restarts"] == 1 and b2["resumed_at"] == 25 and b2["steps_done"] == 30
    for c in range(2):
        np.testing.assert_allclose(np.array(b2["traces"][c]), np.array(a["traces"][c]), rtol=1e-12)
        for ra, rb in zip(a["records"][c], b2["records"][c]):
            assert ra["it"] == rb["it"]
            np.testing.assert_allclose(rb["k"], ra["k"], rtol=1e-12)
            np.testing.assert_allclose(rb["eta_w"], ra["eta_w"], rtol=1e-12)
        assert len(b2["probes"][c]) == len(a["probes"][c])
    with np.load(tmp_path / "b" / "out.npz", allow_pickle=False) as f:
        assert f["replicas/1/ktot"].shape == (6, 30) and int(f["run/restarts"]) == 1
        assert int(f["run/resumed_at_sweep"]) == 25
    assert (tmp_path / "b" / "summary_replicas.txt").exists()


def test_resume_inside_the_warmup_with_training_and_averaged_probes(tmp_path):
    """A run stopped inside the warmup, while the learned proposal still trains (Adam's moments and
    step count in the snapshot) and the step adaptation holds gaps of an unfinished average
    (probe_avg = 2), resumed from its snapshot: the same traces, capacities, Langevin steps and
    learned counts as the run without a stop (FP64 sampler; the network's training is
    reproducible bit for bit)."""
    print("stop file inside the warmup, learned training, averaged probes", flush=True)
    base = dict(init="heuristic", resolution_hours=1.0, nchains=1, replicas=4, mlda_warmup=30, nsamples=6,
                replica_init_sweeps=2, replica_burst_every=4, replica_probe_every=5, replica_probe_len=3,
                replica_probe_tail=1, replica_probe_win=1, replica_probe_avg=2, replica_gate_min=1,
                replica_gate_max=2, replica_gate_min_window=1, replica_gate_max_window=3, replica_gate_win=2,
                synthetic_nodes=10, synthetic_lines=16, snapshots=48, precision="fp64", seed=17,
                learned_proposal=True, learned_train_steps=40, learned_train_every=4, learned_train_more=20,
                learned_components=2)
    a = RP.run_replicas(RP.load_config(dict(base, out=str(tmp_path / "a" / "out.npz"))))
    ck, stop = tmp_path / "ck.npz", tmp_path / "STOP"

    def log(msg):
        if "step 25/" in msg:   # inside the warmup (30 steps), after the fifth probe
            stop.write_text("")

    cfg = RP.load_config(dict(base, checkpoint=str(ck), stop_file=str(stop), out=str(tmp_path / "b" / "out.npz")))
    b1 = RP.run_replicas(cfg, log=log)
    assert b1["status"] == 75 and b1["steps_done"] == 25 and ck.exists()
    with np.load(ck, allow_pickle=False) as f:
        assert f["chains/1/learned/adam"][-1] > 0 and f["chains/1/driver/gaps"].shape[0] == 1
    b2 = RP.run_replicas(cfg)
    assert b2["status"] == 0 and b2["resumed_at"] == 25 and b2["steps_done"] == 36
    np.testing.assert_allclose(np.array(b2["traces"][0]), np.array(a["traces"][0]), rtol=1e-12)
    assert len(b2["probes"][0]) == len(a["probes"][0])
    for pa, pb in zip(a["probes"][0], b2["probes"][0]):
        np.testing.assert_allclose(pb["it"], pa["it"], err_msg="probe step, replica, adapted")
        np.testing.assert_allclose(pb["gap"], pa["gap"], rtol=1e-9, err_msg=f"probe gap at step {pa['it'][0]}")
        np.testing.assert_allclose(pb["eta_after"], pa["eta_after"], rtol=1e-12, err_msg=f"probe eta at {pa['it'][0]}")
    for ra, rb in zip(a["records"][0], b2["records"][0]):
        assert ra["it"] == rb["it"]
        np.testing.assert_allclose(rb["k"], ra["k"], rtol=1e-12, err_msg=f"k at step {ra['it']}")
        np.testing.assert_allclose(rb["eta_w"], ra["eta_w"], rtol=1e-12, err_msg=f"eta_w at step {ra['it']}")
        assert np.array_equal(ra["learned"], rb["learned"])
    assert b2["learned"]["train_steps"] == a["learned"]["train_steps"]


def test_run_replicas_256_nodes_with_the_reference_storage_pairs():
    """The driver on a 256-node / 450-line network with the reference configuration's defaults
    (storage pairs sxpair 0.1 and sxlong 8 on: the hub templates need the 64-lane segments with
    two kink and three balance slots): weather, sweeps, scores, Langevin steps and following run,
    the traces stay finite and the capacities move."""
    print("run_replicas 256 nodes with storage pairs", flush=True)
    cfg = RP.load_config(dict(init="heuristic", resolution_hours=1.0, nchains=1, replicas=4, mlda_warmup=2, nsamples=3,
                              replica_init_sweeps=1, replica_probe_every=0, synthetic_nodes=256,
                              synthetic_lines=450, snapshots=48, precision="fp32", seed=3))
    assert cfg.storage_xpair == 0.1 and cfg.storage_xlong == 8
    res = RP.run_replicas(cfg)
    assert res["status"] == 0 and res["steps_done"] == 5
    tr = np.array(res["traces"][0])
    assert np.all(np.isfinite(tr))
    ks = [r["k"] for r in res["records"][0]]
    assert not np.array_equal(ks[0], ks[-1])


def test_run_replicas_with_ffbs_and_the_reference_replica_configuration():
    """The reference's replica configuration (configs/device40_replica.toml: physical measure, 40
    nodes daily, K = 4 regimes, 10 countries, 4 chains of 4 replicas) with a short schedule and the
    centred FFBS of the regimes at every burst: it runs, regimes move at the bursts, the traces are
    finite."""
    import test_replica_driver as TD
    print("run_replicas reference replica config + ffbs", flush=True)
    cfg = RP.load_config(dict(TD.DEVICE40_REPLICA, mlda_warmup=6, nsamples=6, replica_init_sweeps=2,
                              replica_burst_every=4, replica_probe_every=4, replica_probe_len=3, replica_probe_tail=1,
                              replica_probe_win=1, replica_probe_avg=1, replica_gate_min=1, replica_gate_max=2,
                              replica_gate_min_window=1, replica_gate_max_window=2, replica_gate_win=2,
                              replica_ffbs=True, out=""))
    res = RP.run_replicas(cfg)
    tr = np.array(res["traces"])
    assert tr.shape[:2] == (4, 12) and np.all(np.isfinite(tr))
    assert sum(int(np.sum(r["moved"] > 0)) for r in res["records"][0]) > 0
