"""Build the FP64 CPU oracle (test infrastructure) into ``oracle/liboracle.so``.

    python oracle/build.py

Plain C (gcc, -O2, OpenMP where the sweep uses it). The reference is Julia and has no C or C++
source, so there is no ``oracle/_ref`` build: the restatement is the oracle.
"""
from __future__ import annotations

import os
import pathlib
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
OUT = HERE / "liboracle.so"


def build(verbose: bool = False) -> pathlib.Path:
    srcs = sorted(HERE.glob("*.c"))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-fopenmp", "-ffp-contract=off", "-o", str(tmp),
           *map(str, srcs), "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"gcc failed:\n{p.stderr[-6000:]}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
