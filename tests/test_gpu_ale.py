"""GPU compute_w (fct_compute_w) against the oracle (oracle/fct_numpy.py, itself equal
bit for bit to the C loop restatement), bit for bit: the same per-node (edge, side)
summation order and the same bottom-up cumulative sum."""

import numpy as np
import pytest

from oracle import fct_numpy as O

pytestmark = pytest.mark.gpu


def _t(dev):
    import torch
    return lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("form", [1, 0], ids=["edge_flux", "per_node"])
@pytest.mark.parametrize("which", ["pi", "icosa4_nl70", "icosa4_nl70_deg8", "icosa7"])
def test_compute_w_bitwise(which, form, pi_mesh, pi_case, torch_dev):
    """Both forms of the divergence (FCT_TUNE_W_EDGE_FLUX: the edge terms once per edge
    from the step's edge flux, or per node from the triangles) against the oracle."""
    from fesom_jax_amd import _lib
    from fesom_jax_amd.ale import compute_w
    from fesom_jax_amd.inputs import make_case
    from fesom_jax_amd.meshgen import icosahedral_mesh
    from fesom_jax_amd.tracer_adv import get_plan
    if which == "pi":
        m, c = pi_mesh, pi_case
    elif which == "icosa4_nl70":
        m = icosahedral_mesh(4, 70, seed=41)
        c = make_case(m, seed=42, dt=2700.0)
    elif which == "icosa4_nl70_deg8":
        m = icosahedral_mesh(4, 70, seed=43, degree8=30)     # 8 table slots per node
        c = make_case(m, seed=44, dt=2700.0)
    else:
        m = icosahedral_mesh(7, 70, seed=71)
        c = make_case(m, seed=72, dt=2700.0)
    t = _t(torch_dev)
    plan = get_plan(m, torch_dev)
    _lib.check(plan.lib.fct_plan_set_tuning(plan.ptr, _lib.FCT_TUNE_W_EDGE_FLUX, form), "tuning")
    try:
        got = compute_w(m, t(c["uv"]), t(c["helem"])).cpu().numpy()
    finally:
        plan.lib.fct_plan_set_tuning(plan.ptr, _lib.FCT_TUNE_W_EDGE_FLUX, 1)
    want = O.compute_w(m, c["uv"], c["helem"])
    print(f"[compute_w] {which} form {form}: N={m.nod2D} nl={m.nl} differing="
          f"{int(np.sum(got != want))}", flush=True)
    assert np.array_equal(got, want)


def test_compute_w_numpy_in_numpy_out(pi_mesh, pi_case, torch_dev):
    from fesom_jax_amd.ale import compute_w
    w = compute_w(pi_mesh, pi_case["uv"], pi_case["helem"])
    assert isinstance(w, np.ndarray)
    assert np.array_equal(w, O.compute_w(pi_mesh, pi_case["uv"], pi_case["helem"]))


def test_compute_w_feeds_the_step_bitwise(pi_mesh, pi_case, torch_dev):
    """w from the GPU kernel drives the FCT step exactly like the oracle's w."""
    from fesom_jax_amd.ale import compute_w
    from fesom_jax_amd.tracer_adv import advect_one_fct
    c, t = pi_case, _t(torch_dev)
    w = compute_w(pi_mesh, t(c["uv"]), t(c["helem"]))
    Tn, _ = advect_one_fct(pi_mesh, t(c["uv"]), w, t(c["helem"]), t(c["hnode"]),
                           t(c["hnode_new"]), t(c["T"][0]), t(c["T_old"][0]), dt=c["dt"])
    want = O.advect_one_fct(pi_mesh, c["uv"], O.compute_w(pi_mesh, c["uv"], c["helem"]),
                            c["helem"], c["hnode"], c["hnode_new"], c["T"][0], c["T_old"][0],
                            dt=c["dt"])[0]
    assert np.array_equal(Tn.cpu().numpy(), want)
