"""The committed launch-tuning table and the driver's launch_tuning key (no GPU needed)."""
import json
import re

import pytest

from esmsampling import replicas, tune


def test_tuning_table_entries_are_valid_launch_forms():
    table = json.loads(tune.TABLE.read_text())
    assert table, "the table holds the bench's shape at least"
    assert "N40_L70_T8760_R64_fp32" in table
    for key, e in table.items():
        m = re.fullmatch(r"N(\d+)_L(\d+)_T(\d+)_R(\d+)_fp(32|64)", key)
        assert m, key
        R = int(m.group(4))
        assert e["form"] in ((0, 1, 2, 4) if R % 64 == 0 else (0, 1)), key
        assert e["block"] in (64, 128, 256), key
        assert e["pad"] in ((0,) if R % 64 == 0 else (0, 1)), key
        # the stored choice is the fastest of the timed candidates
        ms = {k: float(v) for k, v in e["ms"].items()}
        assert len(ms) == 12
        assert min(ms, key=ms.get) == f"{e['form']}/{e['block']}/{e['pad']}", key


def test_launch_tuning_key():
    assert replicas.load_config({"init": "heuristic", "sampler": "replicas"}).launch_tuning == "table"
    assert replicas.load_config({"init": "heuristic", "sampler": "replicas", "launch_tuning": "off"}).launch_tuning == "off"
    with pytest.raises(ValueError, match="launch_tuning"):
        replicas.load_config({"init": "heuristic", "sampler": "replicas", "launch_tuning": "live"})


def test_sweep_batch_key():
    assert replicas.load_config({"init": "heuristic", "sampler": "replicas"}).sweep_batch == "graph"
    assert replicas.load_config({"init": "heuristic", "sampler": "replicas", "sweep_batch": "threads"}).sweep_batch == "threads"
    with pytest.raises(ValueError, match="sweep_batch"):
        replicas.load_config({"init": "heuristic", "sampler": "replicas", "sweep_batch": "persistent"})
