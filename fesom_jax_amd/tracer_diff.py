"""Implicit vertical tracer diffusion on the GPU, drop-in for the reference's
``fesom_jax/tracer_diff.py`` (``impl_vert_diff_one`` :68-128, ``impl_vert_diff``
:131-141): same names, arguments, defaults and return values.

One tridiagonal solve per node column (the Thomas recurrence of ``ops.tdma``,
``fesom_jax/ops.py:121-168``, in the reference's operation order), run by the HIP
kernel behind ``fct_impl_vert_diff`` in ``include/fesom_fct.h``. Arrays may be torch
tensors (CUDA/ROCm or CPU), numpy arrays or DLPack producers; results come back in the
kind of ``T``. ``diffuse_tracers`` is the K-tracer form with the step's salinity floor
(``fesom_jax/step.py:452-456``) fused into the same kernel.
"""

from __future__ import annotations

import ctypes
from ctypes import c_void_p

from . import _lib
from .config import DT_DEFAULT
from .tracer_adv import _from_device, _kind_of, _pick_device, _ptr_array, _to_device, _torch, get_plan

__all__ = ["impl_vert_diff_one", "impl_vert_diff", "diffuse_tracers", "S_FLOOR"]

S_FLOOR = 0.5          # fesom_jax/step.py:49


def diffuse_tracers(mesh, Ts, Kv, hnode_new, *, dt: float = DT_DEFAULT, bc_surf=None,
                    sw_3d=None, Z3d=None, floor_tracer=None, floor_value: float = S_FLOOR,
                    out=None):
    """Diffuse K tracers with one ``Kv`` (one matrix per column, solved once per pair of
    tracers). ``bc_surf`` / ``sw_3d``: None or a length-K list whose entries may be None.
    ``floor_tracer``: index of a tracer clamped to ``>= floor_value`` on wet layers
    (``fesom_jax/step.py:456``), None for none. ``out``: optional list of K device
    tensors to write (may be the inputs themselves). Returns a list of device tensors."""
    torch = _torch()
    Ts = list(Ts)
    K = len(Ts)
    if not 1 <= K <= _lib.FCT_VDIFF_MAX_TRACERS:
        raise ValueError(f"need 1..{_lib.FCT_VDIFF_MAX_TRACERS} tracers, got {K}")
    device = _pick_device(*Ts, Kv)
    N, L = int(mesh.nod2D), int(mesh.nl)
    ns = (N, L)
    T_t = [_to_device(T, f"T[{j}]", device, ns) for j, T in enumerate(Ts)]
    kv_t = _to_device(Kv, "Kv", device, ns)
    hnn_t = _to_device(hnode_new, "hnode_new", device, ns)
    z3_t = None if Z3d is None else _to_device(Z3d, "Z3d", device, ns)

    def per_tracer(xs, name, shape):
        if xs is None:
            return None, []
        xs = list(xs)
        if len(xs) != K:
            raise ValueError(f"{name} must have one entry per tracer")
        ts = [None if x is None else _to_device(x, f"{name}[{j}]", device, shape)
              for j, x in enumerate(xs)]
        arr = (c_void_p * K)(*[None if t is None else t.data_ptr() for t in ts])
        return arr, ts

    bc_arr, bc_keep = per_tracer(bc_surf, "bc_surf", (N,))
    sw_arr, sw_keep = per_tracer(sw_3d, "sw_3d", ns)
    if out is None:
        out = [torch.empty(ns, dtype=torch.float64, device=device) for _ in range(K)]
    else:
        out = list(out)
        if len(out) != K:
            raise ValueError("out must have one tensor per tracer")
        for j, o in enumerate(out):
            if not (isinstance(o, torch.Tensor) and o.device == device and
                    o.dtype == torch.float64 and tuple(o.shape) == ns and o.is_contiguous()):
                raise ValueError(f"out[{j}] must be a contiguous float64 tensor of shape {ns} "
                                 f"on {device}")
    if floor_tracer is not None and not 0 <= int(floor_tracer) < K:
        raise ValueError(f"floor_tracer {floor_tracer} outside 0..{K - 1}")
    plan = get_plan(mesh, device)
    Tp, Onp = _ptr_array(T_t), _ptr_array(out)
    P = lambda a: None if a is None else ctypes.cast(a, ctypes.POINTER(c_void_p))
    args = _lib.FctVdiffArgs(
        ntracers=K, floor_tracer=-1 if floor_tracer is None else int(floor_tracer),
        dt=float(dt), floor_value=float(floor_value), Kv=kv_t.data_ptr(),
        hnode_new=hnn_t.data_ptr(), Z3d=None if z3_t is None else z3_t.data_ptr(),
        T=P(Tp), bc_surf=P(bc_arr), sw_3d=P(sw_arr), T_new=P(Onp))
    with torch.cuda.device(device), plan.on(torch.cuda.current_stream(device)):
        stream = torch.cuda.current_stream(device).cuda_stream
        _lib.check(plan.lib.fct_impl_vert_diff(plan.ptr, ctypes.byref(args), c_void_p(stream)),
                   "fct_impl_vert_diff")
    del bc_keep, sw_keep
    return out


def impl_vert_diff_one(mesh, T, Kv, hnode_new, *, dt: float = DT_DEFAULT, bc_surf=None,
                       sw_3d=None, Z3d=None):
    """Per-node implicit vertical diffusion of one tracer; returns the updated tracer
    ``[nod2D, nl]`` (fesom_jax/tracer_diff.py:68-128)."""
    kind = _kind_of(T)
    out = diffuse_tracers(mesh, [T], Kv, hnode_new, dt=dt,
                          bc_surf=None if bc_surf is None else [bc_surf],
                          sw_3d=None if sw_3d is None else [sw_3d], Z3d=Z3d)
    return _from_device(out[0], kind, T)


def impl_vert_diff(mesh, T, S, Kv, hnode_new, *, dt: float = DT_DEFAULT, bc_T=None, bc_S=None,
                   sw_3d=None, Z3d=None):
    """Diffuse T and S with the same ``Kv`` (shortwave on T only); returns
    ``(T_new, S_new)`` (fesom_jax/tracer_diff.py:131-141)."""
    bc = None if bc_T is None and bc_S is None else [bc_T, bc_S]
    sw = None if sw_3d is None else [sw_3d, None]
    out = diffuse_tracers(mesh, [T, S], Kv, hnode_new, dt=dt, bc_surf=bc, sw_3d=sw, Z3d=Z3d)
    return _from_device(out[0], _kind_of(T), T), _from_device(out[1], _kind_of(S), S)
