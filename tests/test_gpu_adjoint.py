"""GPU adjoint (fct_backward) against the CPU reverse-mode oracle (oracle/fct_vjp.py),
which is itself pinned by finite differences (tests/test_oracle_vjp.py). Tolerance:
normwise relative difference <= 1e-11 per cotangent (accumulation orders differ from
the oracle's, so the results agree to rounding, not bit for bit)."""

import numpy as np
import pytest

from conftest import sharp_case
from oracle import fct_numpy as O
from oracle import fct_vjp as V

pytestmark = pytest.mark.gpu

TOL = 1e-11


def _rel(a, b):
    den = max(float(np.abs(b).max()), 1e-300)
    return float(np.abs(a - b).max()) / den


def _oracle(mesh, c, Ts, Tos, gs, **opts):
    out = [V.advect_one_fct_vjp(mesh, c["uv"], c["w_e"], c["helem"], c["hnode"],
                                c["hnode_new"], T, To, g, dt=c["dt"], **opts)
           for T, To, g in zip(Ts, Tos, gs)]
    res = ([o[0] for o in out], [o[1] for o in out], sum(o[2] for o in out),
           sum(o[3] for o in out))
    return res + ((sum(o[4] for o in out),) if opts.get("w_i") is not None else ())


def _gpu(mesh, c, Ts, Tos, gs, dev, **opts):
    import torch
    from fesom_jax_amd.adjoint import advect_fct_vjp
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    opts = {k: (None if v is None else t(v)) for k, v in opts.items()}
    out = advect_fct_vjp(mesh, t(c["uv"]), t(c["w_e"]), t(c["helem"]), t(c["hnode"]),
                         t(c["hnode_new"]), [t(x) for x in Ts], [t(x) for x in Tos],
                         [t(x) for x in gs], dt=c["dt"], **opts)
    torch.cuda.synchronize()
    gT, gTo = out[0], out[1]
    return ([x.cpu().numpy() for x in gT], [x.cpu().numpy() for x in gTo]) + \
        tuple(x.cpu().numpy() for x in out[2:])


def _compare(name, got, want):
    gT, gTo, guv, gw = got[:4]
    wT, wTo, wuv, ww = want[:4]
    assert len(got) == len(want)
    for j in range(len(gT)):
        assert _rel(gT[j], wT[j]) <= TOL, (name, "g_T", j, _rel(gT[j], wT[j]))
        assert _rel(gTo[j], wTo[j]) <= TOL, (name, "g_T_old", j, _rel(gTo[j], wTo[j]))
    assert _rel(guv, wuv) <= TOL, (name, "g_uv", _rel(guv, wuv))
    assert _rel(gw, ww) <= TOL, (name, "g_w", _rel(gw, ww))
    if len(got) > 4:
        assert _rel(got[4], want[4]) <= TOL, (name, "g_w_i", _rel(got[4], want[4]))


def _split(mesh, c, seed):
    """An implicit split w_i of both signs and a moving-level stencil (Z3d, zbar3)."""
    rng = np.random.default_rng(seed)
    wm = mesh.node_iface_mask
    w_i = np.where(wm, 0.3 * np.abs(c["w_e"]).max() * rng.standard_normal(wm.shape), 0.0)
    zb3 = mesh.zbar_3d_n * (1.0 + 0.01 * rng.random(mesh.zbar_3d_n.shape))
    Z3d = np.concatenate([0.5 * (zb3[:, :-1] + zb3[:, 1:]), zb3[:, -1:]], axis=1)
    return w_i, Z3d, zb3


def _cot(mesh, K, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal((mesh.nod2D, mesh.nl)) for _ in range(K)]


def test_adjoint_pi_two_tracers(pi_mesh, pi_case, torch_dev):
    c = pi_case
    Ts = c["T"]
    Tos = [0.995 * T for T in Ts]
    gs = _cot(pi_mesh, 2, 1)
    _compare("pi", _gpu(pi_mesh, c, Ts, Tos, gs, torch_dev), _oracle(pi_mesh, c, Ts, Tos, gs))


def test_adjoint_pi_limiter_active(pi_mesh, pi_case, torch_dev):
    c = sharp_case(pi_mesh, pi_case, 2.0)
    gs = _cot(pi_mesh, 1, 2)
    Ts, Tos = c["T"][:1], c["T_old"][:1]
    _, _, p = O.advect_one_fct(pi_mesh, c["uv"], c["w_e"], c["helem"], c["hnode"],
                               c["hnode_new"], Ts[0], Tos[0], dt=c["dt"], return_parts=True)
    assert (p["plus_fac"][pi_mesh.node_layer_mask] < 1).mean() > 0.02
    _compare("pi sharp", _gpu(pi_mesh, c, Ts, Tos, gs, torch_dev),
             _oracle(pi_mesh, c, Ts, Tos, gs))


def test_adjoint_icosa_nl70(torch_dev):
    from fesom_jax_amd.inputs import make_case
    from fesom_jax_amd.meshgen import icosahedral_mesh
    m = icosahedral_mesh(3, 70, seed=51)
    c = make_case(m, seed=52, dt=2700.0)
    Ts = c["T"]
    Tos = [0.99 * T + 0.01 * T[::-1] for T in Ts]
    gs = _cot(m, 2, 3)
    _compare("icosa3 nl70", _gpu(m, c, Ts, Tos, gs, torch_dev), _oracle(m, c, Ts, Tos, gs))


def test_adjoint_is_deterministic(pi_mesh, pi_case, torch_dev):
    c = pi_case
    gs = _cot(pi_mesh, 2, 4)
    a = _gpu(pi_mesh, c, c["T"], c["T_old"], gs, torch_dev)
    b = _gpu(pi_mesh, c, c["T"], c["T_old"], gs, torch_dev)
    for x, y in zip(a[0] + a[1] + [a[2], a[3]], b[0] + b[1] + [b[2], b[3]]):
        assert np.array_equal(x, y)


def test_autograd_function_matches_vjp(pi_mesh, pi_case, torch_dev):
    import torch
    from fesom_jax_amd.adjoint import fct_step
    c = pi_case
    t = lambda a, rg=False: torch.from_numpy(np.ascontiguousarray(a)).to(torch_dev).requires_grad_(rg)
    uv, w = t(c["uv"], True), t(c["w_e"], True)
    T = [t(x, True) for x in c["T"]]
    To = [t(0.99 * x, True) for x in c["T"]]
    out = fct_step(pi_mesh, uv, w, t(c["helem"]), t(c["hnode"]), t(c["hnode_new"]), T, To,
                   dt=c["dt"])
    gs = _cot(pi_mesh, 2, 5)
    loss = sum((o * t(g)).sum() for o, g in zip(out, gs))
    grads = torch.autograd.grad(loss, [*T, *To, uv, w])
    want = _oracle(pi_mesh, c, c["T"], [0.99 * x for x in c["T"]], gs)
    for j in range(2):
        assert _rel(grads[j].cpu().numpy(), want[0][j]) <= TOL
        assert _rel(grads[2 + j].cpu().numpy(), want[1][j]) <= TOL
    assert _rel(grads[4].cpu().numpy(), want[2]) <= TOL
    assert _rel(grads[5].cpu().numpy(), want[3]) <= TOL


def test_rollout_gradient_matches_oracle_chain(torch_dev):
    """Four chained steps: torch autograd through FctStep against the oracle's
    reverse chain (T_old <- T, T <- T_new), loss on the surface level."""
    import torch
    from fesom_jax_amd.adjoint import rollout
    from fesom_jax_amd.inputs import make_case
    from fesom_jax_amd.meshgen import icosahedral_mesh
    m = icosahedral_mesh(3, 48, seed=61)
    c = make_case(m, seed=62, dt=2700.0)
    nsteps = 4
    rng = np.random.default_rng(7)
    target = [T + 0.2 * rng.standard_normal(T.shape) * m.node_layer_mask for T in c["T"]]
    surf = m.node_layer_mask[:, 0]
    # oracle: forward trajectory, then the reverse chain
    traj = [(list(c["T"]), list(c["T_old"]))]
    for _ in range(nsteps):
        T, To = traj[-1]
        Tn = [O.advect_one_fct(m, c["uv"], c["w_e"], c["helem"], c["hnode"], c["hnode_new"],
                               T[j], To[j], dt=c["dt"])[0] for j in range(2)]
        traj.append((Tn, T))
    gT = [np.zeros_like(x) for x in c["T"]]
    for j in range(2):
        gT[j][:, 0] = np.where(surf, 2.0 * (traj[-1][0][j][:, 0] - target[j][:, 0]), 0.0)
    gTo = [np.zeros_like(x) for x in c["T"]]
    guv = np.zeros_like(c["uv"])
    for s in range(nsteps - 1, -1, -1):
        T, To = traj[s]
        res = [V.advect_one_fct_vjp(m, c["uv"], c["w_e"], c["helem"], c["hnode"],
                                    c["hnode_new"], T[j], To[j], gT[j], dt=c["dt"])
               for j in range(2)]
        guv = guv + sum(r[2] for r in res)
        # T_s feeds step s (as T) and step s+1 (as T_old)
        gT, gTo = [r[0] + gTo[j] for j, r in enumerate(res)], [r[1] for r in res]
    want_T0 = [gT[j] for j in range(2)]
    want_To0 = gTo
    # GPU: autograd through the rollout
    t = lambda a, rg=False: torch.from_numpy(np.ascontiguousarray(a)).to(torch_dev).requires_grad_(rg)
    uv = t(c["uv"], True)
    T0 = [t(x, True) for x in c["T"]]
    To0 = [t(x, True) for x in c["T_old"]]
    fin = rollout(m, uv, t(c["w_e"]), t(c["helem"]), t(c["hnode"]), t(c["hnode_new"]), T0,
                  To0, nsteps, dt=c["dt"])
    sm = t(surf.astype(np.float64))
    loss = sum((((f[:, 0] - t(tg[:, 0])) ** 2) * sm).sum() for f, tg in zip(fin, target))
    grads = torch.autograd.grad(loss, [*T0, *To0, uv])
    for j in range(2):
        assert _rel(grads[j].cpu().numpy(), want_T0[j]) <= 1e-10
        assert _rel(grads[2 + j].cpu().numpy(), want_To0[j]) <= 1e-10
    assert _rel(grads[4].cpu().numpy(), guv) <= 1e-10


def test_backward_rejects_unsupported(pi_mesh, pi_case, torch_dev):
    import torch
    from fesom_jax_amd.adjoint import advect_fct_vjp
    c = pi_case
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(torch_dev)
    with pytest.raises(ValueError):
        advect_fct_vjp(pi_mesh, t(c["uv"]), t(c["w_e"]), t(c["helem"]), t(c["hnode"]),
                       t(c["hnode_new"]), [t(c["T"][0])], [t(c["T_old"][0])], [], dt=c["dt"])


@pytest.mark.parametrize("kind", ["w_i", "Z3d", "both"])
def test_adjoint_split_and_zstar(kind, torch_dev):
    """The implicit split (the Thomas solve reversed, w_e + w_i in adf_v, g_w_i) and the
    moving-level QR4C stencil, smooth and limiter-active."""
    from fesom_jax_amd.inputs import make_case
    from fesom_jax_amd.meshgen import icosahedral_mesh
    m = icosahedral_mesh(3, 48, seed=41)
    c0 = make_case(m, seed=42, dt=2700.0)
    for sharp in (False, True):
        c = sharp_case(m, c0, 2.0) if sharp else c0
        w_i, Z3d, zb3 = _split(m, c, 7)
        opts = {"w_i": dict(w_i=w_i), "Z3d": dict(Z3d=Z3d, zbar3=zb3),
                "both": dict(w_i=w_i, Z3d=Z3d, zbar3=zb3)}[kind]
        Ts, Tos = c["T"][:2], [0.99 * T + 0.01 * T[::-1] for T in c["T"][:2]]
        if sharp:
            Tos = c["T_old"][:2]
        gs = _cot(m, len(Ts), 3)
        print(f"[adjoint] {kind} sharp={sharp}", flush=True)
        _compare(f"{kind} sharp={sharp}", _gpu(m, c, Ts, Tos, gs, torch_dev, **opts),
                 _oracle(m, c, Ts, Tos, gs, **opts))


def test_adjoint_split_pi(pi_mesh, pi_case, torch_dev):
    c = pi_case
    w_i, Z3d, zb3 = _split(pi_mesh, c, 17)
    Ts, Tos, gs = c["T"][:1], [0.995 * c["T"][0]], _cot(pi_mesh, 1, 18)
    opts = dict(w_i=w_i, Z3d=Z3d, zbar3=zb3)
    _compare("pi split", _gpu(pi_mesh, c, Ts, Tos, gs, torch_dev, **opts),
             _oracle(pi_mesh, c, Ts, Tos, gs, **opts))
