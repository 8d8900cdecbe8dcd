"""Records, outputs and snapshots of the replica driver (CPU): the datasets and layouts of
``write_replica_outputs`` (docs/system/data.md §3: one column per step), the text summary, the
ragged driver buffers of a snapshot, and the snapshot checks of ``read_replica_snapshot``
(src/replicas.jl:1133-1167: driver, version, seed)."""
import numpy as np
import pytest

from esmsampling import api
from esmsampling import records as rio
from esmsampling import replicas as RP
from esmsampling import synth


def _recs(rng, steps, M, N, nf=0):
    out = []
    for it in range(1, steps + 1):
        parts = rng.uniform(1, 2, (M, 8))
        parts[:, 0] = 7.0
        r = rio.replica_step_record(rng.uniform(0, 100, (6, N)), parts, np.full(M, 1.0 / M), rng.random(M) < 0.5,
                                    np.arange(M), np.zeros(M), rng.normal(size=6), rng.uniform(1, 2, 6), np.full(6, 5e-4),
                                    3, 0.5, 0.4, [10, 4])
        r.update(it=it, k=np.zeros((6, N)), nclip=3)
        if nf:
            r["features"] = rng.normal(size=M * nf)
        out.append(r)
    return out


def test_outputs_layout_and_summary(tmp_path):
    rng = np.random.default_rng(1)
    cfg = RP.load_config({"init": "heuristic", "replicas": 3, "nchains": 2, "mlda_warmup": 2, "nsamples": 3})
    ctl = RP.ReplicaCtl.from_config(cfg)
    s = synth.make_system(N=5, L=6, T=24, dt=1.0, seed=2)
    names = api.replica_feature_names(s)
    nf = len(names)
    assert names[:4] == ["load_GW", "cf_solar", "cf_onwind", "cf_offwind"] and names[-1] == "regime_switches"
    assert nf == 3 + 5 + s.K
    recs = [_recs(rng, 5, 3, s.N, nf) for _ in range(2)]
    pr = {"gap": np.ones(6), "rs_eq": np.ones(6), "rs_in": np.ones(6), "fk": np.ones(6), "op_in": 1.0, "op_eq": 2.0,
          "shed_in": 0.0, "shed_eq": 0.0}
    probes = [[rio.probe_record(4, 1, True, pr, np.ones(6), 2 * np.ones(6))], []]
    path = tmp_path / "run.npz"
    rio.write_replica_outputs(str(path), cfg, s, ctl, recs, probes, restarts=1, resumed_at=3)
    with np.load(path, allow_pickle=False) as f:
        d = {k: f[k] for k in f.files}
    assert list(d["replicas/carriers"]) == list(synth.CARRIERS) and int(d["replicas/M"]) == 3
    assert d["replicas/1/ktot"].shape == (6, 5) and d["replicas/1/op"].shape == (3, 5)
    assert d["replicas/2/capex"].shape == (1, 5) and np.all(d["replicas/2/capex"] == 7.0)
    assert d["replicas/1/time"].shape == (2, 5) and d["replicas/1/learned"].shape == (2, 5)
    assert np.array_equal(d["replicas/1/pfcost"][:, 2], recs[0][2]["pfcost"])     # column = step
    assert d["replicas/1/probe/it"].shape == (3, 1) and list(d["replicas/1/probe/it"][:, 0]) == [4.0, 2.0, 1.0]
    assert "replicas/2/probe/it" not in d
    assert int(d["run/restarts"]) == 1 and int(d["run/resumed_at_sweep"]) == 3
    assert d["weather/transition"].shape == s.Ptr.shape
    # features stacked replica by replica, one column per step (src/replicas.jl:818)
    assert list(d["replicas/feature_names"]) == names
    assert d["replicas/2/features"].shape == (3 * nf, 5)
    assert np.array_equal(d["replicas/2/features"][:, 4], recs[1][4]["features"])
    sp = rio.summary_path(str(path))
    rio.write_replica_summary(sp, ctl, recs, probes, ["low", "high"], 2, False, 12.0, feature_names=names)
    txt = open(sp).read()
    assert "## chain 1 (low), steps 3–5" in txt and "## chain 2 (high)" in txt and "probes" in txt
    assert "replica weather features" in txt
    # the summary's feature means over the production steps 3–5 and all replicas
    X = np.concatenate([r["features"].reshape(3, nf) for r in recs[0][2:]])
    line = [q for q in txt.splitlines() if q.strip().startswith("resload_peak_GW")][0]
    assert abs(float(line.split()[1]) - X[:, names.index("resload_peak_GW")].mean()) < 1e-5 * (1 + abs(X[:, 6].mean()))


def test_ragged_buffers_round_trip():
    rng = np.random.default_rng(3)
    lists = [[rng.normal(size=(6, 4)) for _ in range(n)] for n in (0, 3, 1)]
    vals, lens, shape = rio._ragged(lists)
    back = rio._unragged(vals, lens, shape)
    assert [len(x) for x in back] == [0, 3, 1]
    for a, b in zip(lists, back):
        for u, w in zip(a, b):
            assert np.array_equal(u, w)


def test_rng_state_round_trip():
    g = np.random.Generator(np.random.Philox(key=[3, 4000]))
    g.standard_normal(7)
    h = np.random.Generator(np.random.Philox(key=[1, 1]))
    h.bit_generator.state = rio.rng_state_from_json(rio.rng_state_json(g.bit_generator.state))
    assert np.array_equal(g.standard_normal(5), h.standard_normal(5))


def test_snapshot_checks(tmp_path):
    cfg = RP.load_config({"init": "heuristic", "seed": 4, "nchains": 2})
    assert rio.read_snapshot(str(tmp_path / "none.npz"), cfg) is None
    p = tmp_path / "ck.npz"
    base = {"version": np.array(1), "driver": np.array("replicas"), "config/seed": np.array(4), "nchains": np.array(2)}
    rio._atomic_savez(str(p), dict(base, driver=np.array("chains")))
    with pytest.raises(ValueError, match="not a snapshot of the replica driver"):
        rio.read_snapshot(str(p), cfg)
    rio._atomic_savez(str(p), dict(base, version=np.array(2)))
    with pytest.raises(ValueError, match="format version"):
        rio.read_snapshot(str(p), cfg)
    rio._atomic_savez(str(p), dict(base, **{"config/seed": np.array(5)}))
    with pytest.raises(ValueError, match="seed"):
        rio.read_snapshot(str(p), cfg)
    rio._atomic_savez(str(p), base)
    assert rio.read_snapshot(str(p), cfg)["nchains"] == 2
    assert not any(q.name.startswith(".tmp_") for q in tmp_path.iterdir())   # atomic writes leave no temporaries


def test_snapshot_config_keys():
    c = RP.load_config({"init": "heuristic", "stop_file": "/tmp/x", "checkpoint_minutes": 5.0})
    assert c.checkpoint == "" and c.stop_file == "/tmp/x" and c.checkpoint_minutes == 5.0
    assert RP.load_config({"init": "heuristic"}).checkpoint_minutes == 30.0


def test_records_copy_the_drivers_arrays():
    """The driver adapts its η and other buffers in place: a step's or a probe's record keeps the
    values of its own step (an alias would show the last step's values in every column)."""
    M = 3
    eta, wts, act = np.full(6, 5e-4), np.full(M, 1.0 / M), np.ones(M, dtype=bool)
    rec = rio.replica_step_record(np.ones((6, 2)), np.ones((M, 8)), wts, act, np.arange(M), np.zeros(M),
                                  np.zeros(6), np.ones(6), eta, 0, 0.1, 0.1, [1, 1])
    pr = dict(gap=np.zeros(6), rs_eq=np.zeros(6), rs_in=np.zeros(6), fk=np.ones(6), op_in=1.0, op_eq=1.0,
              shed_in=0.0, shed_eq=0.0)
    before = eta.copy()
    eta *= 1.2
    prec = rio.probe_record(5, 0, True, pr, before, eta)
    eta *= 1.2
    wts[:] = 0.0
    act[:] = False
    np.testing.assert_array_equal(rec["eta_w"], np.full(6, 5e-4))
    np.testing.assert_array_equal(rec["weights"], np.full(M, 1.0 / M))
    assert rec["active"].all()
    np.testing.assert_array_equal(prec["eta_after"], before * 1.2)
