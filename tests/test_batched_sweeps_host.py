"""The replica driver's graph mode on the host side (no GPU): the barrier hands every chain's
sweep to one batched call and each chain its own statistics; a chain that fails releases the
others with an error instead of leaving them waiting."""
import threading

import numpy as np
import pytest

from esmsampling import api, replicas


class FakeBatch:
    """Stands in for api.Batch: records the calls and returns statistics that name the chain."""

    calls = []

    def __init__(self, devs):
        self.devs = devs

    def sweeps(self, nsweeps, keys, counters, maxnats):
        FakeBatch.calls.append((list(keys), list(counters), maxnats))
        return np.stack([np.full((15, 5), 100 * c + counters[c], dtype=np.int64) for c in range(len(self.devs))])

    def close(self):
        pass


def _run(fn, n):
    out, errs = [None] * n, [None] * n

    def work(c):
        try:
            out[c] = fn(c)
        except BaseException as e:   # noqa: BLE001 - the test inspects it
            errs[c] = e

    ts = [threading.Thread(target=work, args=(c,)) for c in range(n)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=30)
    assert not any(t.is_alive() for t in ts), "a chain is still waiting"
    return out, errs


def test_each_chain_gets_its_own_sweep(monkeypatch):
    monkeypatch.setattr(api, "Batch", FakeBatch)
    FakeBatch.calls = []
    devs = [object() for _ in range(3)]
    bs = replicas.BatchedSweeps(devs)
    out, errs = _run(lambda c: bs.sweep(devs[c], (7, c), 10 + c, 40.0), 3)
    assert errs == [None] * 3
    assert FakeBatch.calls == [([(7, 0), (7, 1), (7, 2)], [10, 11, 12], 40.0)]
    for c in range(3):
        assert np.all(out[c] == 100 * c + 10 + c)


def test_a_failing_chain_releases_the_others(monkeypatch):
    monkeypatch.setattr(api, "Batch", FakeBatch)
    devs = [object() for _ in range(3)]
    bs = replicas.BatchedSweeps(devs)

    def chain(c):
        if c == 1:
            bs.abort()            # what run_replicas does when a chain's step raises
            raise RuntimeError("chain 1 failed")
        return bs.sweep(devs[c], (1, c), c, 40.0)

    out, errs = _run(chain, 3)
    assert isinstance(errs[1], RuntimeError)
    assert all(isinstance(errs[c], threading.BrokenBarrierError) for c in (0, 2))


def test_one_maxnats_per_batched_sweep(monkeypatch):
    monkeypatch.setattr(api, "Batch", FakeBatch)
    devs = [object() for _ in range(2)]
    bs = replicas.BatchedSweeps(devs)
    out, errs = _run(lambda c: bs.sweep(devs[c], (1, c), c, 40.0 + c), 2)
    # the barrier's action runs in the last chain to arrive: that chain gets the error, the other
    # one a broken barrier
    assert sorted(type(e).__name__ for e in errs) == ["BrokenBarrierError", "ValueError"]
