"""The reference's stage functions (fesom_jax/tracer_adv.py:59-529), each run alone on the
GPU through fct_stage, against the oracle's stage of the same name on the same inputs,
bit for bit. Inputs are the oracle's own intermediates of a full step (smooth and
limiter-active states), so every stage sees the values it sees inside the step."""

import numpy as np
import pytest

from conftest import sharp_case
from oracle import fct_numpy as O

pytestmark = pytest.mark.gpu


def _cases(pi_mesh, pi_case, icosa4_mesh, icosa4_case):
    from fesom_jax_amd.inputs import make_case
    from fesom_jax_amd.meshgen import icosahedral_mesh
    m70 = icosahedral_mesh(4, 70, seed=23, degree8=6)
    c70 = make_case(m70, seed=24, dt=2700.0)
    return [("pi", pi_mesh, pi_case), ("pi sharp", pi_mesh, sharp_case(pi_mesh, pi_case, 2.0)),
            ("icosa4", icosa4_mesh, icosa4_case),
            ("icosa4 nl70 sharp", m70, sharp_case(m70, c70, 2.0))]


def _eq(name, got, want):
    g = got.cpu().numpy() if hasattr(got, "cpu") else np.asarray(got)
    assert g.shape == want.shape, (name, g.shape, want.shape)
    ndiff = int(np.sum(g != want))
    print(f"[stages] {name}: differing entries {ndiff}/{g.size}", flush=True)
    assert ndiff == 0, (name, float(np.max(np.abs(g - want))))


@pytest.fixture(scope="module")
def cases(pi_mesh, pi_case, icosa4_mesh, icosa4_case):
    out = []
    for name, m, c in _cases(pi_mesh, pi_case, icosa4_mesh, icosa4_case):
        T, To = c["T"][0], c["T_old"][0]
        _, _, parts = O.advect_one_fct(m, c["uv"], c["w_e"], c["helem"], c["hnode"],
                                       c["hnode_new"], T, To, dt=c["dt"], return_parts=True)
        out.append((name, m, c, T, To, parts))
    return out


def test_flux_stages(cases, torch_dev):
    import torch
    from fesom_jax_amd import tracer_adv as G
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(torch_dev)
    for name, m, c, T, To, p in cases:
        ttf = O.tracer_ab(T, To)[0]
        g_ttf, same = G.tracer_ab(t(T), t(To))
        _eq(f"{name} tracer_ab", g_ttf, ttf)
        _eq(f"{name} edge_vflux", G.edge_vflux(m, t(c["uv"]), t(c["helem"])),
            O.edge_vflux(m, c["uv"], c["helem"]))
        _eq(f"{name} adv_flux_hor", G.adv_flux_hor(m, t(c["uv"]), t(c["helem"]), t(T)),
            O.adv_flux_hor(m, c["uv"], c["helem"], T))
        _eq(f"{name} adv_flux_ver", G.adv_flux_ver(m, t(c["w_e"]), t(T)),
            O.adv_flux_ver(m, c["w_e"], T))
        _eq(f"{name} compute_fct_lo",
            G.compute_fct_lo(m, t(p["flux_h_lo"]), t(p["flux_v_lo"]), t(T), t(c["hnode"]),
                             t(c["hnode_new"]), dt=c["dt"]),
            O.compute_fct_lo(m, p["flux_h_lo"], p["flux_v_lo"], T, c["hnode"], c["hnode_new"],
                             dt=c["dt"]))
        _eq(f"{name} tracer_gradient_elements", G.tracer_gradient_elements(m, t(T)),
            O.tracer_gradient_elements(m, T))
        _eq(f"{name} fill_up_dn_grad", G.fill_up_dn_grad(m, t(p["tr_xy"])),
            O.fill_up_dn_grad(m, p["tr_xy"]))
        _eq(f"{name} adv_flux_hor_ho",
            G.adv_flux_hor_ho(m, t(c["uv"]), t(c["helem"]), t(ttf), t(p["eud"])),
            O.adv_flux_hor_ho(m, c["uv"], c["helem"], ttf, p["eud"]))
        _eq(f"{name} adv_flux_ver_ho", G.adv_flux_ver_ho(m, t(c["w_e"]), t(ttf)),
            O.adv_flux_ver_ho(m, c["w_e"], ttf))


def test_zstar_vertical_flux(cases, torch_dev):
    import torch
    from fesom_jax_amd import tracer_adv as G
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(torch_dev)
    rng = np.random.default_rng(3)
    for name, m, c, T, To, p in cases[:1] + cases[-1:]:
        s = rng.uniform(0.9, 1.1, (m.nod2D, 1))
        zbar3 = np.asarray(m.zbar)[None, :] * s
        Z3d = np.concatenate([0.5 * (zbar3[:, :-1] + zbar3[:, 1:]), zbar3[:, -1:]], axis=1)
        ttf = O.tracer_ab(T, To)[0]
        _eq(f"{name} adv_flux_ver_ho Z3d", G.adv_flux_ver_ho(m, t(c["w_e"]), t(ttf), t(Z3d),
                                                             t(zbar3)),
            O.adv_flux_ver_ho(m, c["w_e"], ttf, Z3d=Z3d, zbar3=zbar3))


def test_limiter_and_update_stages(cases, torch_dev):
    import torch
    from fesom_jax_amd import tracer_adv as G
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(torch_dev)
    for name, m, c, T, To, p in cases:
        args = (T, p["LO"], p["adf_h"], p["adf_v"], c["hnode_new"])
        want = O.zalesak_limit(m, *args, dt=c["dt"])
        got = G.zalesak_limit(m, *[t(x) for x in args], dt=c["dt"])
        _eq(f"{name} zalesak_limit adf_h", got[0], want[0])
        _eq(f"{name} zalesak_limit adf_v", got[1], want[1])
        # the exch route: P+/P- through the hook, scaled by an exact power of two on the
        # way so that the hook's result is what the factors are computed from
        seen = []

        def exch(f, kind, seen=seen):
            seen.append(kind)
            return f * 0.5 if isinstance(f, np.ndarray) else f * 0.5
        want_x = O.zalesak_limit(m, *args, dt=c["dt"], exch=exch)
        got_x = G.zalesak_limit(m, *[t(x) for x in args], dt=c["dt"], exch=exch)
        assert seen == ["nod", "nod", "nod", "nod"]
        _eq(f"{name} zalesak_limit exch adf_h", got_x[0], want_x[0])
        _eq(f"{name} zalesak_limit exch adf_v", got_x[1], want_x[1])
        fargs = (T, p["LO"], p["adf_h_lim"], p["adf_v_lim"], c["hnode"], c["hnode_new"])
        d_want = O.flux2dtracer_fct(m, *fargs, dt=c["dt"])
        _eq(f"{name} flux2dtracer_fct", G.flux2dtracer_fct(m, *[t(x) for x in fargs], dt=c["dt"]),
            d_want)
        _eq(f"{name} ale_reconstruct",
            G.ale_reconstruct(m, t(T), t(d_want), t(c["hnode"]), t(c["hnode_new"])),
            O.ale_reconstruct(m, T, d_want, c["hnode"], c["hnode_new"]))
        _eq(f"{name} flux2dtracer",
            G.flux2dtracer(m, t(p["flux_h_lo"]), t(p["flux_v_lo"]), dt=c["dt"]),
            O.flux2dtracer(m, p["flux_h_lo"], p["flux_v_lo"], dt=c["dt"]))


def test_vertical_implicit_stage(cases, torch_dev):
    import torch
    from fesom_jax_amd import tracer_adv as G
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(torch_dev)
    rng = np.random.default_rng(6)
    for name, m, c, T, To, p in cases:
        w_i = np.where(m.node_iface_mask, 0.3 * c["w_e"] * rng.random(m.node_iface_mask.shape),
                       0.0)
        _eq(f"{name} adv_tra_vert_impl",
            G.adv_tra_vert_impl(m, t(w_i), t(p["LO"]), t(c["hnode_new"]), dt=c["dt"]),
            O.adv_tra_vert_impl(m, w_i, p["LO"], c["hnode_new"], dt=c["dt"]))


def test_stage_argument_errors(pi_mesh, torch_dev):
    import torch
    from fesom_jax_amd import tracer_adv as G
    N, L = pi_mesh.nod2D, pi_mesh.nl
    x = torch.zeros((N, L), dtype=torch.float64, device=torch_dev)
    with pytest.raises(ValueError):
        G.adv_flux_ver(pi_mesh, x[:, :-1], x)
    with pytest.raises(ValueError):
        G.adv_flux_ver_ho(pi_mesh, x, x, Z3d=x)
    with pytest.raises(ValueError):
        G.adv_flux_ver(pi_mesh, x.float(), x)
