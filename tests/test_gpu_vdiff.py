"""GPU implicit vertical diffusion (fct_impl_vert_diff) against the oracle
(oracle/vert_diff.py), bit for bit: the kernel runs the reference's Thomas recurrence
per column in the same operation order, compiled without FP contraction."""

import numpy as np
import pytest

from oracle import vert_diff as VD

pytestmark = pytest.mark.gpu


def _kv(mesh, seed):
    rng = np.random.default_rng(seed)
    return np.where(mesh.node_iface_mask, rng.uniform(1e-5, 3e-2, (mesh.nod2D, mesh.nl)), 0.0)


def _forcing(mesh, seed):
    rng = np.random.default_rng(seed)
    bc = rng.standard_normal(mesh.nod2D) * 0.3
    sw = np.where(mesh.node_iface_mask, 1e-4 * np.exp(np.asarray(mesh.zbar)[None, :] / 30.0), 0.0)
    return bc, sw * (1.0 + 0.1 * rng.standard_normal(sw.shape))


def _t(dev):
    import torch
    return lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _eq(got, want, name):
    g = got.cpu().numpy()
    assert np.array_equal(g, want), (name, float(np.abs(g - want).max()))


@pytest.fixture(scope="module")
def icosa7():
    from fesom_jax_amd.inputs import make_case
    from fesom_jax_amd.meshgen import icosahedral_mesh
    m = icosahedral_mesh(7, 70, seed=71)
    return m, make_case(m, seed=72, dt=2700.0)


def test_one_tracer_pi(pi_mesh, pi_case, torch_dev):
    from fesom_jax_amd.tracer_diff import impl_vert_diff_one
    m, c, t = pi_mesh, pi_case, _t(torch_dev)
    Kv = _kv(m, 1)
    got = impl_vert_diff_one(m, t(c["T"][0]), t(Kv), t(c["hnode_new"]), dt=c["dt"])
    _eq(got, VD.impl_vert_diff_one(m, c["T"][0], Kv, c["hnode_new"], dt=c["dt"]), "pi T")


def test_two_tracers_with_forcing_pi(pi_mesh, pi_case, torch_dev):
    from fesom_jax_amd.tracer_diff import impl_vert_diff
    m, c, t = pi_mesh, pi_case, _t(torch_dev)
    Kv = _kv(m, 2)
    bc, sw = _forcing(m, 3)
    bcS = -0.5 * bc
    T2, S2 = impl_vert_diff(m, t(c["T"][0]), t(c["T"][1]), t(Kv), t(c["hnode_new"]),
                            dt=c["dt"], bc_T=t(bc), bc_S=t(bcS), sw_3d=t(sw))
    wT, wS = VD.impl_vert_diff(m, c["T"][0], c["T"][1], Kv, c["hnode_new"], dt=c["dt"],
                               bc_T=bc, bc_S=bcS, sw_3d=sw)
    _eq(T2, wT, "pi T forced")
    _eq(S2, wS, "pi S forced")


def test_live_depths_icosa4_nl70(torch_dev):
    from fesom_jax_amd.inputs import make_case
    from fesom_jax_amd.meshgen import icosahedral_mesh
    from fesom_jax_amd.tracer_diff import impl_vert_diff
    m = icosahedral_mesh(4, 70, seed=41)
    c = make_case(m, seed=42, dt=2700.0)
    t = _t(torch_dev)
    rng = np.random.default_rng(4)
    Zs = np.concatenate([m.Z, m.Z[-1:]])
    Z3d = Zs[None, :] * (1.0 + 0.05 * rng.uniform(-1, 1, (m.nod2D, 1)))
    Kv = _kv(m, 5)
    T2, S2 = impl_vert_diff(m, t(c["T"][0]), t(c["T"][1]), t(Kv), t(c["hnode_new"]),
                            dt=c["dt"], Z3d=t(Z3d))
    wT, wS = VD.impl_vert_diff(m, c["T"][0], c["T"][1], Kv, c["hnode_new"], dt=c["dt"], Z3d=Z3d)
    _eq(T2, wT, "icosa4 Z3d T")
    _eq(S2, wS, "icosa4 Z3d S")


def test_many_blocks_per_workgroup_icosa7(icosa7, torch_dev):
    """163842 columns: more column blocks than resident workgroups, so each workgroup
    loops and reuses its scratch slot."""
    from fesom_jax_amd.tracer_diff import diffuse_tracers
    m, c = icosa7
    t = _t(torch_dev)
    Kv = _kv(m, 7)
    bc, sw = _forcing(m, 8)
    out = diffuse_tracers(m, [t(c["T"][0]), t(c["T"][1])], t(Kv), t(c["hnode_new"]),
                          dt=c["dt"], bc_surf=[t(bc), None], sw_3d=[t(sw), None],
                          floor_tracer=1, floor_value=35.0)
    wT, wS = VD.impl_vert_diff(m, c["T"][0], c["T"][1], Kv, c["hnode_new"], dt=c["dt"],
                               bc_T=bc, sw_3d=sw)
    _eq(out[0], wT, "icosa7 T")
    _eq(out[1], VD.salinity_floor(m, wS, 35.0), "icosa7 S floored")


def test_three_tracers_and_in_place(pi_mesh, pi_case, torch_dev):
    from fesom_jax_amd.tracer_diff import diffuse_tracers
    m, c, t = pi_mesh, pi_case, _t(torch_dev)
    Kv = _kv(m, 9)
    Ts = [c["T"][0], c["T"][1], 0.5 * c["T"][0] + 3.0]
    bc, _ = _forcing(m, 10)
    want = [VD.impl_vert_diff_one(m, T, Kv, c["hnode_new"], dt=c["dt"],
                                  bc_surf=bc if j == 2 else None) for j, T in enumerate(Ts)]
    want[2] = VD.salinity_floor(m, want[2], 20.0)
    dev = [t(T) for T in Ts]
    out = diffuse_tracers(m, dev, t(Kv), t(c["hnode_new"]), dt=c["dt"],
                          bc_surf=[None, None, t(bc)], floor_tracer=2, floor_value=20.0,
                          out=dev)
    for j in range(3):
        assert out[j].data_ptr() == dev[j].data_ptr()
        _eq(dev[j], want[j], f"in place tracer {j}")


def test_reference_properties_on_gpu(icosa7, torch_dev):
    """The reference's property test (fesom_jax/tests/test_tracers.py:214-235) at
    163842 columns: constant-in-z stays, column integral conserved, gradient smoothed."""
    from fesom_jax_amd.tracer_diff import impl_vert_diff_one
    m, c = icosa7
    t = _t(torch_dev)
    wet = m.node_layer_mask
    Kv_big = np.where(m.node_iface_mask, 1.0, 0.0)
    Tc = np.where(wet, 20.0, 0.0)
    Tc2 = impl_vert_diff_one(m, t(Tc), t(Kv_big), t(c["hnode_new"]), dt=c["dt"]).cpu().numpy()
    assert np.max(np.abs(Tc2[wet] - 20.0)) < 1e-12
    k = np.arange(m.nl)[None, :]
    Tg = np.where(wet, 10.0 + 0.5 * k, 0.0)
    Tg2 = impl_vert_diff_one(m, t(Tg), t(Kv_big), t(c["hnode_new"]), dt=c["dt"]).cpu().numpy()
    av, hn = np.asarray(m.areasvol), np.asarray(c["hnode_new"])
    i0 = np.sum(av * hn * Tg, axis=1)
    i1 = np.sum(av * hn * Tg2, axis=1)
    assert np.max(np.abs(i1 - i0) / np.maximum(np.abs(i0), 1e-30)) < 1e-12
    rng0 = np.where(wet, Tg, -1e30).max(1) - np.where(wet, Tg, 1e30).min(1)
    rng1 = np.where(wet, Tg2, -1e30).max(1) - np.where(wet, Tg2, 1e30).min(1)
    assert rng1.mean() < rng0.mean()


def test_rejects_bad_arguments(pi_mesh, pi_case, torch_dev):
    from fesom_jax_amd.tracer_diff import diffuse_tracers
    m, c, t = pi_mesh, pi_case, _t(torch_dev)
    Kv = t(_kv(m, 11))
    with pytest.raises(ValueError):
        diffuse_tracers(m, [], Kv, t(c["hnode_new"]))
    with pytest.raises(ValueError):
        diffuse_tracers(m, [t(c["T"][0])], Kv, t(c["hnode_new"]), floor_tracer=1)
    with pytest.raises(ValueError):
        diffuse_tracers(m, [t(c["T"][0])], Kv, t(c["hnode_new"]), bc_surf=[None, None])
    # an output that is another tracer's input is refused by the library
    T0, T1 = t(c["T"][0]), t(c["T"][1])
    with pytest.raises(ValueError):
        diffuse_tracers(m, [T0, T1], Kv, t(c["hnode_new"]), out=[T1, T0])
    # a non-contiguous output is refused rather than written through a copy
    with pytest.raises(ValueError):
        diffuse_tracers(m, [T0], Kv, t(c["hnode_new"]),
                        out=[t(np.asfortranarray(c["T"][0])).t().contiguous().t()])
