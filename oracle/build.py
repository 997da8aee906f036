"""Build recipe of the C loop oracle (test infrastructure only): compiles
oracle/fct_loops.c with gcc, without floating-point contraction, into
oracle/_build/libfct_loops.so. Run by ``__graft_entry__.build()``; tests build it on
first use if it is missing."""

from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "fct_loops.c"
OUT = HERE / "_build" / "libfct_loops.so"
# the same loops with their OpenMP directives on (the multi-core CPU baseline); each
# node-level still sums its contributions in the serial order, so the bits are the same
OUT_OMP = HERE / "_build" / "libfct_loops_omp.so"


def build(verbose: bool = False) -> Path:
    OUT.parent.mkdir(exist_ok=True)
    base = ["gcc", "-O2", "-std=c99", "-fPIC", "-shared", "-ffp-contract=off", "-Wall"]
    for out, extra in ((OUT, ["-Wno-unknown-pragmas"]), (OUT_OMP, ["-fopenmp"])):
        cmd = base + extra + ["-o", str(out), str(SRC), "-lm"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return OUT


def ensure(openmp: bool = False) -> Path:
    out = OUT_OMP if openmp else OUT
    if not out.exists() or out.stat().st_mtime < SRC.stat().st_mtime:
        build()
    return out


if __name__ == "__main__":
    print(build(verbose=True))
