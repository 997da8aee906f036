"""Shared fixtures. GPU tests are marked ``gpu``; everything else runs on the CPU."""

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"
PI_MESH_DIR = GOLDEN / "mesh_pi"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


@pytest.fixture(scope="session")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no ROCm GPU visible")
    return torch.device("cuda", 0)


@pytest.fixture(scope="session")
def pi_mesh():
    from fesom_jax_amd.mesh import load_mesh
    return load_mesh(PI_MESH_DIR)


@pytest.fixture(scope="session")
def pi_case(pi_mesh):
    from fesom_jax_amd.inputs import make_case
    return make_case(pi_mesh, seed=0, dt=2700.0)


@pytest.fixture(scope="session")
def icosa4_mesh():
    from fesom_jax_amd.meshgen import icosahedral_mesh
    return icosahedral_mesh(4, 48, seed=3)


@pytest.fixture(scope="session")
def icosa4_case(icosa4_mesh):
    from fesom_jax_amd.inputs import make_case
    return make_case(icosa4_mesh, seed=4, dt=2700.0, tracers=("T", "S", "T"))


def rel_err(a, b, mask=None):
    """max |a-b| / max |b| over the (optional) mask."""
    a = np.asarray(a); b = np.asarray(b)
    if mask is not None:
        m = np.broadcast_to(mask, a.shape)
        a, b = a[m], b[m]
    scale = float(np.max(np.abs(b))) if b.size else 1.0
    return float(np.max(np.abs(a - b))) / (scale if scale > 0 else 1.0)


def sharp_case(mesh, case, factor):
    """The reference's limiter-active recipe (fesom_jax/tests/test_tracers.py:578-588):
    a sharp tracer 10 + 5 sin(0.7 n) cos(1.3 k), T_old = 10, flow scaled by ``factor``."""
    k = np.arange(mesh.nl)[None, :]
    m = mesh.node_layer_mask
    sharp = np.where(m, 10.0 + 5.0 * np.sin(np.arange(mesh.nod2D)[:, None] * 0.7)
                     * np.cos(k * 1.3), 0.0)
    told = np.where(m, 10.0, 0.0)
    return dict(case, uv=case["uv"] * factor, w_e=case["w_e"] * factor, T=[sharp],
                T_old=[told])
