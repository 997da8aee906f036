"""GPU sea-ice FCT (fct_ice_fct_solve) against the oracle (oracle/ice_fct.py, itself
equal bit for bit to the C element loops), bit for bit: per node, the elements in
increasing id, three-vertex sums ((x0 + x1) + x2), the element factor sqrt(vol / 2e8)
correctly rounded on the host."""

import copy
import ctypes
import math

import numpy as np
import pytest

from oracle import ice_fct as I

pytestmark = pytest.mark.gpu


def _t(dev):
    import torch
    return lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _gpu(cfg, mesh, c, dev):
    from fesom_jax_amd.ice_adv import fct_solve
    t = _t(dev)
    out = fct_solve(cfg, mesh, t(c["a_ice"]), t(c["m_ice"]), t(c["m_snow"]), t(c["u_ice"]),
                    t(c["v_ice"]))
    return [o.cpu().numpy() for o in out]


def _want(cfg, mesh, c):
    return I.fct_solve(cfg, mesh, c["a_ice"], c["m_ice"], c["m_snow"], c["u_ice"], c["v_ice"])


def _shelf(mesh, frac=0.1, seed=5):
    rng = np.random.default_rng(seed)
    m = copy.copy(mesh)
    ul = np.array(mesh.ulevels, copy=True)
    ul[rng.random(mesh.elem2D) < frac] = 2
    open_node = np.zeros(mesh.nod2D, bool)
    for q in range(3):
        open_node[mesh.elem_nodes[ul <= 1, q]] = True
    uln = np.where(open_node, 1, 2).astype(np.asarray(mesh.ulevels_nod2D).dtype)
    ulmax = np.zeros(mesh.nod2D, np.int64)
    for q in range(3):
        np.maximum.at(ulmax, mesh.elem_nodes[:, q], ul)
    object.__setattr__(m, "ulevels", ul)
    object.__setattr__(m, "ulevels_nod2D", uln)
    object.__setattr__(m, "ulevels_nod2D_max",
                       ulmax.astype(np.asarray(mesh.ulevels_nod2D_max).dtype))
    return m


@pytest.mark.parametrize("dt", [500.0, 2.0e4])
def test_ice_bitwise_pi(dt, pi_mesh, torch_dev):
    from fesom_jax_amd.inputs import make_ice_case
    cfg = I.IceConfig(ice_dt=dt)
    c = make_ice_case(pi_mesh, seed=3, speed=0.5)
    for name, g, w in zip(("a", "m", "ms"), _gpu(cfg, pi_mesh, c, torch_dev),
                          _want(cfg, pi_mesh, c)):
        print(f"[ice] pi dt={dt} {name}: differing={int(np.sum(g != w))}", flush=True)
        assert np.array_equal(g, w), name


def test_ice_bitwise_icosa_with_shelf(torch_dev):
    from fesom_jax_amd.inputs import make_ice_case
    from fesom_jax_amd.meshgen import icosahedral_mesh
    m = _shelf(icosahedral_mesh(4, 24, seed=43))
    cfg = I.IceConfig(ice_dt=3.0e4, ice_gamma_fct=0.25, ice_diff=5.0)
    c = make_ice_case(m, seed=8, speed=0.5)
    for g, w in zip(_gpu(cfg, m, c, torch_dev), _want(cfg, m, c)):
        assert np.array_equal(g, w)


def test_ice_large_mesh_bitwise_and_mass(torch_dev):
    """163,842 nodes: bitwise, and the area-weighted mass of each tracer conserved."""
    from fesom_jax_amd.inputs import make_ice_case
    from fesom_jax_amd.meshgen import icosahedral_mesh
    m = icosahedral_mesh(7, 24, seed=73)
    cfg = I.IceConfig(ice_dt=4.0e3)
    c = make_ice_case(m, seed=9, speed=0.5)
    got = _gpu(cfg, m, c, torch_dev)
    ar = np.asarray(m.area)[:, 0]
    for g, w, x0 in zip(got, _want(cfg, m, c), (c["a_ice"], c["m_ice"], c["m_snow"])):
        assert np.array_equal(g, w)
        m0, m1 = math.fsum(ar * x0), math.fsum(ar * g)
        assert abs(m1 - m0) <= 1e-13 * abs(m0)


def test_ice_numpy_in_numpy_out_and_errors(pi_mesh, torch_dev):
    from fesom_jax_amd import _lib
    from fesom_jax_amd.ice_adv import IceConfig, fct_solve
    from fesom_jax_amd.inputs import make_ice_case
    from fesom_jax_amd.tracer_adv import get_plan
    c = make_ice_case(pi_mesh, seed=4)
    cfg = IceConfig(ice_dt=1.0e4)
    out = fct_solve(cfg, pi_mesh, c["a_ice"], c["m_ice"], c["m_snow"], c["u_ice"], c["v_ice"])
    assert all(isinstance(o, np.ndarray) for o in out)
    want = _want(I.IceConfig(ice_dt=1.0e4), pi_mesh, c)
    assert all(np.array_equal(o, w) for o, w in zip(out, want))
    with pytest.raises(NotImplementedError):
        fct_solve(cfg, pi_mesh, c["a_ice"], c["m_ice"], c["m_snow"], c["u_ice"], c["v_ice"],
                  exch=lambda f, k: f)
    # the C ABI refuses an output that aliases an input
    import torch
    t = _t(torch_dev)
    ins = [t(c[k]) for k in ("a_ice", "m_ice", "m_snow", "u_ice", "v_ice")]
    plan = get_plan(pi_mesh, torch_dev)
    args = _lib.FctIceArgs(ice_dt=1e4, ice_diff=10.0, ice_gamma_fct=0.5,
                           a_ice=ins[0].data_ptr(), m_ice=ins[1].data_ptr(),
                           m_snow=ins[2].data_ptr(), u_ice=ins[3].data_ptr(),
                           v_ice=ins[4].data_ptr(), a_new=ins[0].data_ptr(),
                           m_new=torch.empty_like(ins[0]).data_ptr(),
                           ms_new=torch.empty_like(ins[0]).data_ptr())
    rc = plan.lib.fct_ice_fct_solve(plan.ptr, ctypes.byref(args), None)
    assert rc == _lib.FCT_E_INVALID
