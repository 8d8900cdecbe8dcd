"""Fill the launch-tuning table of the operation sweep on the GPU.

    python -m esmsampling.tune [--out esmsampling/tuning.json] [--shape N,L,T,R,prec ...] [--warm 4]

For every workload shape (nodes, lines, hours, replicas, precision 32 or 64) it builds the seeded
synthetic system, draws the weather, takes the heuristic state a few operation sweeps away from
its start (walk lengths change as the state leaves the box corners) and runs ``api.autotune``:
each launch form of the random families (``sweep_launch``) with each workgroup size
(``sweep_block``) timed on that state, which is restored afterwards. The choice of each shape is
stored in the JSON table under ``api._shape_key``; ``api.autotune(R, cache=table)`` later applies
it without timing (the bench and the replica driver do).
"""
from __future__ import annotations

import argparse
import pathlib
import time

TABLE = pathlib.Path(__file__).resolve().parent / "tuning.json"

# the bench's step (40 n, 8760 h, 64 r), the 128-replica and 2920 h shapes of the learned leg and
# the driver, and the 256-node shapes of the uneven batch
DEFAULT_SHAPES = [(40, 70, 8760, 64, 32), (40, 70, 8760, 128, 32), (40, 70, 2920, 16, 32), (40, 70, 2920, 96, 32),
                  (256, 450, 2920, 64, 32), (256, 450, 2920, 32, 32)]


def tune_shape(N, L, T, R, prec, table, warm=4, seed=1):
    from . import FP32, FP64, api, synth
    s = synth.make_system(N=N, L=L, T=T, dt=1.0, seed=seed)
    d = api.Replicas(s, R, FP64 if prec == 64 else FP32)
    try:
        api.sample_weather(d, key=(1, 0))
        d.upload(v=synth.heuristic_state(s, d.download()[2]), k=s.k0)
        for c in range(warm):
            api.operation_sweep(d, 1, key=(2, 2), counter=c)
        t0 = time.perf_counter()
        res = api.autotune(d, nsweeps=2, key=(3, 3), counter=0, cache=str(table), retune=True)
        ms = {f"{f}/{b}/{p}": round(t, 2) for (f, b, p), t in res["ms"].items()}
        print(f"{api._shape_key(d)}: form {res['form']} block {res['block']} pad {res['pad']} "
              f"({time.perf_counter() - t0:.1f} s) {ms}", flush=True)
        return res
    finally:
        d.close()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=str(TABLE))
    ap.add_argument("--shape", action="append", default=[], help="N,L,T,R,prec (repeatable)")
    ap.add_argument("--warm", type=int, default=4, help="operation sweeps before timing")
    a = ap.parse_args(argv)
    shapes = [tuple(int(x) for x in sh.split(",")) for sh in a.shape] or DEFAULT_SHAPES
    out = pathlib.Path(a.out)
    for sh in shapes:
        tune_shape(*sh, table=out, warm=a.warm)


if __name__ == "__main__":
    main()
