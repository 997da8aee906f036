"""bench.py's contract pieces that need no GPU: the one-GPU refusal, the work and byte
model the metric is quoted in, and the committed traffic figure it reports."""

import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_work_and_byte_model_count_wet_entries(pi_mesh):
    from fesom_jax_amd.perfmodel import adjoint_bytes, algorithmic_bytes, updates_per_step
    m = pi_mesh
    # one node-level update per tracer and wet node layer, counted from the masks
    wet_layers = int(np.asarray(m.node_layer_mask).sum())
    assert updates_per_step(m, 3) == 3 * wet_layers
    # the step's bytes grow by exactly 3 wet node layers of float64 per extra tracer
    # (T and T_old read, T_new written), and the backward's by 5 (T, T_old, g_T_new
    # read; g_T, g_T_old written)
    d1 = algorithmic_bytes(m, 3) - algorithmic_bytes(m, 2)
    assert d1 == 8 * 3 * wet_layers
    assert adjoint_bytes(m, 3) - adjoint_bytes(m, 2) == 8 * 5 * wet_layers
    assert adjoint_bytes(m, 2) > algorithmic_bytes(m, 2)


def test_committed_traffic_figure_parses():
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench", ROOT / "bench.py")
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)

    class Args:
        level, seed = 9, 2024

    class M:
        nl = 70
    t = bench.measured_traffic(Args, M, 2)
    assert isinstance(t, int) and 12_000_000_000 < t < 1_000_000_000_000
    Args.level = 8
    assert bench.measured_traffic(Args, M, 2) is None
