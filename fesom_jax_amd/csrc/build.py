"""Build libfesom_fct.so (gfx950) in-tree with hipcc.

Flags: -ffp-contract=off keeps every FP64 expression in the operation order of the
reference's jnp expression (no fused multiply-add), which is what makes the kernels
reproduce the CPU restatement bit for bit; packed FP32 instructions are disabled.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
PKG = HERE.parent
ROOT = PKG.parent
LIB = PKG / "libfesom_fct.so"
SOURCES = ["fct_api.hip", "fct_phased.hip", "fct_fused.hip", "fct_adjoint.hip", "fct_vdiff.hip", "fct_ale.hip",
           "fct_ice.hip", "fct_locality.hip"]
ARCH = os.environ.get("FESOM_FCT_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the FCT library needs the ROCm toolchain")


def build(verbose: bool = False) -> Path:
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wall", "-Wno-unused-function",
           f"-I{ROOT / 'include'}", f"-I{HERE}", "-o", str(LIB) + ".tmp"]
    cmd += [str(HERE / s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode})")
    os.replace(str(LIB) + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
