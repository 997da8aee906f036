"""Sea-ice FCT advection on the GPU, drop-in for the reference's
``fesom_jax.ice_adv.fct_solve`` (``fesom_jax/ice_adv.py:177-194``): same name,
arguments and return ``(a_new, m_new, ms_new)``. The HIP kernels behind
``fct_ice_fct_solve`` (``include/fesom_fct.h``) gather, per surface node, over the
elements around it in increasing element id: Taylor-Galerkin right-hand side,
consistent-mass low- and high-order solves, element antidiffusive fluxes, Zalesak ratios
and limiting. ``cfg`` is any object with ``ice_dt``, ``ice_diff`` and ``ice_gamma_fct``
(the reference's ``IceConfig`` works as is). Arrays may be torch tensors, numpy arrays or
DLPack producers; results come back in the kind of ``a_ice``."""

from __future__ import annotations

import ctypes
from ctypes import c_void_p
from dataclasses import dataclass

from . import _lib
from .tracer_adv import _from_device, _kind_of, _pick_device, _to_device, _torch, get_plan

__all__ = ["IceConfig", "fct_solve"]


@dataclass
class IceConfig:
    """The settings the ice FCT reads, with the reference's defaults
    (``fesom_jax/ice.py:58-62``)."""
    ice_gamma_fct: float = 0.5
    ice_diff: float = 10.0
    ice_dt: float = 500.0


def fct_solve(cfg, mesh, a_ice, m_ice, m_snow, u_ice, v_ice, exch=None):
    """Advance the three ice tracers one FCT step; returns ``(a_new, m_new, ms_new)``
    (before ``cut_off``, as the reference)."""
    if exch is not None:
        raise NotImplementedError("fct_solve: the exch hook (sharded runs) is not supported; "
                                  "pass exch=None")
    torch = _torch()
    device = _pick_device(a_ice, m_ice, u_ice)
    N = int(mesh.nod2D)
    ins = [_to_device(x, name, device, (N,)) for x, name in
           ((a_ice, "a_ice"), (m_ice, "m_ice"), (m_snow, "m_snow"), (u_ice, "u_ice"),
            (v_ice, "v_ice"))]
    outs = [torch.empty((N,), dtype=torch.float64, device=device) for _ in range(3)]
    args = _lib.FctIceArgs(ice_dt=float(cfg.ice_dt), ice_diff=float(cfg.ice_diff),
                           ice_gamma_fct=float(cfg.ice_gamma_fct),
                           a_ice=ins[0].data_ptr(), m_ice=ins[1].data_ptr(),
                           m_snow=ins[2].data_ptr(), u_ice=ins[3].data_ptr(),
                           v_ice=ins[4].data_ptr(), a_new=outs[0].data_ptr(),
                           m_new=outs[1].data_ptr(), ms_new=outs[2].data_ptr())
    plan = get_plan(mesh, device)
    with torch.cuda.device(device), plan.on(torch.cuda.current_stream(device)):
        stream = torch.cuda.current_stream(device).cuda_stream
        _lib.check(plan.lib.fct_ice_fct_solve(plan.ptr, ctypes.byref(args), c_void_p(stream)),
                   "fct_ice_fct_solve")
    like = (a_ice, m_ice, m_snow)
    return tuple(_from_device(o, _kind_of(x), x) for o, x in zip(outs, like))
