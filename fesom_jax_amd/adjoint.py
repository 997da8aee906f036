"""Reverse mode of the FCT step on the GPU.

The reference differentiates ``advect_one_fct`` (``fesom_jax/tracer_adv.py:532-578``)
with ``jax.grad`` (no custom VJP, ``docs/LIMITER_GRADIENTS.md:45-75``). Here the
backward is hand-written HIP (``fct_backward`` in ``include/fesom_fct.h``): it recomputes
the step's forward intermediates and runs the chain rule back with jax.grad's
conventions for max/min ties and ``abs``. ``FctStep`` wraps forward and backward in a
``torch.autograd.Function`` so that rollouts of many steps differentiate with torch's
autograd, storing only each step's inputs.
"""

from __future__ import annotations

import ctypes
from ctypes import c_void_p

from . import _lib
from .config import DT_DEFAULT
from .tracer_adv import _pick_device, _ptr_array, _to_device, _torch, advect_fct, get_plan


def advect_fct_vjp(mesh, uv, w_e, helem, hnode, hnode_new, Ts, T_olds, g_T_news, *,
                   dt=DT_DEFAULT, need_uv=True, need_w=True, plan=None, Z3d=None,
                   zbar3=None, w_i=None):
    """Cotangents of one K-tracer FCT step. Returns ``(g_Ts, g_T_olds, g_uv, g_w_e)``
    (torch tensors on the device; ``g_uv`` / ``g_w_e`` summed over the tracers, or
    None when not requested), and ``g_w_i`` as a fifth entry when the implicit split
    ``w_i`` is given. ``Z3d`` / ``zbar3`` select the moving-level QR4C stencil (they,
    ``helem`` and the thicknesses are constants). ``plan`` (an ``FctPlan`` of ``mesh``
    on the tensors' device) holds the adjoint workspace; by default the mesh's cached
    plan does."""
    torch = _torch()
    device = _pick_device(Ts[0], uv)
    N, E, L = int(mesh.nod2D), int(mesh.elem2D), int(mesh.nl)
    ns, es = (N, L), (E, L)
    if (Z3d is None) != (zbar3 is None):
        raise ValueError("Z3d and zbar3 must be given together")
    uv_t = _to_device(uv, "uv", device, (E, L, 2))
    w_t = _to_device(w_e, "w_e", device, ns)
    wi_t = None if w_i is None else _to_device(w_i, "w_i", device, ns)
    z_t = None if Z3d is None else _to_device(Z3d, "Z3d", device, ns)
    zb_t = None if zbar3 is None else _to_device(zbar3, "zbar3", device, ns)
    he_t = _to_device(helem, "helem", device, es)
    hn_t = _to_device(hnode, "hnode", device, ns)
    hnn_t = _to_device(hnode_new, "hnode_new", device, ns)
    T_t = [_to_device(x, f"T[{j}]", device, ns) for j, x in enumerate(Ts)]
    To_t = [_to_device(x, f"T_old[{j}]", device, ns) for j, x in enumerate(T_olds)]
    g_t = [_to_device(x, f"g_T_new[{j}]", device, ns) for j, x in enumerate(g_T_news)]
    if not (len(T_t) == len(To_t) == len(g_t)):
        raise ValueError("Ts, T_olds and g_T_news must have the same length")
    f64 = dict(dtype=torch.float64, device=device)
    gT = [torch.empty(ns, **f64) for _ in T_t]
    gTo = [torch.empty(ns, **f64) for _ in T_t]
    g_uv = torch.empty((E, L, 2), **f64) if need_uv else None
    g_w = torch.empty(ns, **f64) if need_w else None
    g_wi = None if wi_t is None else torch.empty(ns, **f64)
    if plan is None:
        plan = get_plan(mesh, device)
    elif plan.device != device or plan.nod2D != N or plan.nl != L:
        raise ValueError("plan does not belong to this mesh and device")
    keep = [_ptr_array(T_t), _ptr_array(To_t), _ptr_array(g_t), _ptr_array(gT), _ptr_array(gTo)]
    P = lambda a: ctypes.cast(a, ctypes.POINTER(c_void_p))
    fwd = _lib.FctStepArgs(ntracers=len(T_t), dt=float(dt), uv=uv_t.data_ptr(),
                           w_e=w_t.data_ptr(), helem=he_t.data_ptr(), hnode=hn_t.data_ptr(),
                           hnode_new=hnn_t.data_ptr(), T=P(keep[0]), T_old=P(keep[1]),
                           T_new=None, w_i=None if wi_t is None else wi_t.data_ptr(),
                           Z3d=None if z_t is None else z_t.data_ptr(),
                           zbar3=None if zb_t is None else zb_t.data_ptr())
    adj = _lib.FctAdjointArgs(g_T_new=P(keep[2]), g_T=P(keep[3]), g_T_old=P(keep[4]),
                              g_uv=None if g_uv is None else g_uv.data_ptr(),
                              g_w_e=None if g_w is None else g_w.data_ptr(),
                              g_w_i=None if g_wi is None else g_wi.data_ptr())
    with torch.cuda.device(device), plan.on(torch.cuda.current_stream(device)):
        s = torch.cuda.current_stream(device).cuda_stream
        _lib.check(plan.lib.fct_backward(plan.ptr, ctypes.byref(fwd), ctypes.byref(adj),
                                         c_void_p(s)), "fct_backward")
    if g_wi is not None:
        return gT, gTo, g_uv, g_w, g_wi
    return gT, gTo, g_uv, g_w


def advect_one_fct_vjp(mesh, uv, w_e, helem, hnode, hnode_new, T, T_old, g_T_new, *,
                       dt=DT_DEFAULT, Z3d=None, zbar3=None, w_i=None):
    """Single-tracer form: ``(g_T, g_T_old, g_uv, g_w_e)`` (+ ``g_w_i`` with ``w_i``)."""
    out = advect_fct_vjp(mesh, uv, w_e, helem, hnode, hnode_new, [T], [T_old], [g_T_new],
                         dt=dt, Z3d=Z3d, zbar3=zbar3, w_i=w_i)
    return (out[0][0], out[1][0]) + tuple(out[2:])


def _fct_step_function():
    torch = _torch()

    class FctStep(torch.autograd.Function):
        """T_new = FCT(T, T_old; uv, w_e) for K tracers. Differentiable in every T,
        T_old, uv and w_e; helem, hnode, hnode_new are treated as constants. Saves only
        the inputs: the backward recomputes the step."""

        @staticmethod
        def forward(ctx, mesh, dt, K, uv, w_e, helem, hnode, hnode_new, *tracers):
            Ts, T_olds = list(tracers[:K]), list(tracers[K:])
            outs = advect_fct(mesh, uv, w_e, helem, hnode, hnode_new, Ts, T_olds, dt=dt)
            ctx.mesh, ctx.dt, ctx.K = mesh, dt, K
            ctx.save_for_backward(uv, w_e, helem, hnode, hnode_new, *tracers)
            return tuple(o[0] for o in outs)

        @staticmethod
        def backward(ctx, *g_outs):
            uv, w_e, helem, hnode, hnode_new, *tracers = ctx.saved_tensors
            K = ctx.K
            g = [torch.zeros_like(tracers[j]) if go is None else go.contiguous()
                 for j, go in enumerate(g_outs)]
            need_uv, need_w = ctx.needs_input_grad[3], ctx.needs_input_grad[4]
            gT, gTo, g_uv, g_w = advect_fct_vjp(
                ctx.mesh, uv, w_e, helem, hnode, hnode_new, tracers[:K], tracers[K:], g,
                dt=ctx.dt, need_uv=need_uv, need_w=need_w)
            return (None, None, None, g_uv, g_w, None, None, None, *gT, *gTo)

    return FctStep


_FCT_STEP = None


def fct_step(mesh, uv, w_e, helem, hnode, hnode_new, Ts, T_olds, *, dt=DT_DEFAULT):
    """Differentiable K-tracer FCT step (torch tensors in, list of T_new out)."""
    global _FCT_STEP
    if _FCT_STEP is None:
        _FCT_STEP = _fct_step_function()
    K = len(Ts)
    return list(_FCT_STEP.apply(mesh, float(dt), K, uv, w_e, helem, hnode, hnode_new,
                                *Ts, *T_olds))


def rollout(mesh, uv, w_e, helem, hnode, hnode_new, Ts, T_olds, nsteps, *, dt=DT_DEFAULT):
    """``nsteps`` chained steps (T_old <- T, T <- T_new, ``tracer_adv.py:549``);
    differentiable through ``fct_step``. Returns the final tracers."""
    T, To = list(Ts), list(T_olds)
    for _ in range(nsteps):
        Tn = fct_step(mesh, uv, w_e, helem, hnode, hnode_new, T, To, dt=dt)
        To, T = T, Tn
    return T
