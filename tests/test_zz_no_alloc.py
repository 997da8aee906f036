"""fct_impl_vert_diff_static (the vertical diffusion that allocates nothing, its scratch a
fixed device array) against fct_impl_vert_diff (plan scratch), bit for bit, through the
C ABI: two calls in a row on the same GPU, the second on a smaller mesh and with fewer
tracers, so the second call reuses the array the first one left behind."""

import ctypes
from ctypes import c_void_p

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(entry, plan, dev, T, Kv, hnn, dt, *, bc=None, sw=None, floor=None):
    import torch
    from fesom_jax_amd import _lib
    K = len(T)
    out = [torch.empty_like(T[0]) for _ in range(K)]
    Tp = (c_void_p * K)(*[t.data_ptr() for t in T])
    Op = (c_void_p * K)(*[o.data_ptr() for o in out])
    bcp = None if bc is None else (c_void_p * K)(*[None if b is None else b.data_ptr() for b in bc])
    swp = None if sw is None else (c_void_p * K)(*[None if w is None else w.data_ptr() for w in sw])
    P = lambda a: None if a is None else ctypes.cast(a, ctypes.POINTER(c_void_p))
    args = _lib.FctVdiffArgs(ntracers=K, floor_tracer=-1 if floor is None else floor[0],
                             dt=dt, floor_value=0.0 if floor is None else floor[1],
                             Kv=Kv.data_ptr(), hnode_new=hnn.data_ptr(), Z3d=None, T=P(Tp),
                             bc_surf=P(bcp), sw_3d=P(swp), T_new=P(Op))
    stream = torch.cuda.current_stream(dev).cuda_stream
    _lib.check(getattr(plan.lib, entry)(plan.ptr, ctypes.byref(args), c_void_p(stream)), entry)
    torch.cuda.synchronize(dev)
    return [o.cpu().numpy() for o in out]


def _case(level, nl, K, seed, dev):
    import torch
    from fesom_jax_amd.inputs import make_case
    from fesom_jax_amd.meshgen import icosahedral_mesh
    from fesom_jax_amd.tracer_adv import get_plan
    m = icosahedral_mesh(level, nl, seed=seed)
    c = make_case(m, seed=seed + 1, dt=2700.0, tracers=("T", "S")[:K])
    rng = np.random.default_rng(seed + 2)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    Kv = np.where(m.node_iface_mask, rng.uniform(1e-5, 3e-2, (m.nod2D, m.nl)), 0.0)
    bc = [t(rng.standard_normal(m.nod2D) * 0.3) for _ in range(K)]
    sw = [t(np.where(m.node_iface_mask, 1e-4 * np.exp(np.asarray(m.zbar)[None, :] / 30.0), 0.0))]
    sw += [None] * (K - 1)
    return (m, get_plan(m, dev), [t(x) for x in c["T"][:K]], t(Kv), t(c["hnode_new"]), bc, sw)


def test_static_scratch_matches_allocating_call_twice(torch_dev):
    # first call: icosa-6, nl = 70, two tracers with surface fluxes, shortwave and floor
    m, plan, T, Kv, hnn, bc, sw = _case(6, 70, 2, 40, torch_dev)
    print(f"[no-alloc] call 1: icosa-6 N={m.nod2D} nl={m.nl} K=2", flush=True)
    kw = dict(bc=bc, sw=sw, floor=(1, 34.9))
    got = _call("fct_impl_vert_diff_static", plan, torch_dev, T, Kv, hnn, 2700.0, **kw)
    want = _call("fct_impl_vert_diff", plan, torch_dev, T, Kv, hnn, 2700.0, **kw)
    for j in range(2):
        assert np.array_equal(got[j], want[j]), ("call 1", j)
    # second call: a smaller mesh (icosa-4, nl = 48) and one tracer
    m2, plan2, T2, Kv2, hnn2, bc2, sw2 = _case(4, 48, 1, 50, torch_dev)
    print(f"[no-alloc] call 2: icosa-4 N={m2.nod2D} nl={m2.nl} K=1", flush=True)
    got2 = _call("fct_impl_vert_diff_static", plan2, torch_dev, T2, Kv2, hnn2, 1800.0,
                 bc=bc2, sw=sw2)
    want2 = _call("fct_impl_vert_diff", plan2, torch_dev, T2, Kv2, hnn2, 1800.0, bc=bc2, sw=sw2)
    assert np.array_equal(got2[0], want2[0]), "call 2"
    # the diffusion changed the tracer (the comparison is not between two copies)
    assert not np.array_equal(got2[0], T2[0].cpu().numpy())


def test_static_scratch_argument_errors(torch_dev):
    from fesom_jax_amd import _lib
    m, plan, T, Kv, hnn, bc, sw = _case(3, 48, 1, 60, torch_dev)
    args = _lib.FctVdiffArgs(ntracers=0, floor_tracer=-1, dt=1.0, floor_value=0.0,
                             Kv=Kv.data_ptr(), hnode_new=hnn.data_ptr(), Z3d=None, T=None,
                             bc_surf=None, sw_3d=None, T_new=None)
    rc = plan.lib.fct_impl_vert_diff_static(plan.ptr, ctypes.byref(args), None)
    assert rc == _lib.FCT_E_INVALID
    assert b"ntracers" in plan.lib.fct_last_error()
