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


def build(verbose: bool = False) -> Path:
    OUT.parent.mkdir(exist_ok=True)
    cmd = ["gcc", "-O2", "-std=c99", "-fPIC", "-shared", "-ffp-contract=off", "-Wall",
           "-o", str(OUT), str(SRC), "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return OUT


def ensure() -> Path:
    if not OUT.exists() or OUT.stat().st_mtime < SRC.stat().st_mtime:
        build()
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
