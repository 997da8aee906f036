"""Vertical velocity from continuity on the GPU, drop-in for the reference's
``fesom_jax.ale.compute_w`` (``fesom_jax/ale.py:309-365``): same name, arguments and
return. The HIP kernel behind ``fct_compute_w`` (``include/fesom_fct.h``) sums each
node's edge divergence in the reference's (edge, side) order, integrates it from the
bottom in the original C loop's order and divides by the interface area. Arrays may be
torch tensors, numpy arrays or DLPack producers; the result comes back in the kind of
``uv``."""

from __future__ import annotations

from ctypes import c_void_p

from . import _lib
from .tracer_adv import _from_device, _kind_of, _pick_device, _to_device, _torch, get_plan

__all__ = ["compute_w"]


def compute_w(mesh, uv, helem):
    """``w`` ``[nod2D, nl]`` at the interfaces (m/s), zero outside each node's
    interfaces ``[nzmin, nzmax]`` (fesom_jax/ale.py:309-365)."""
    torch = _torch()
    kind = _kind_of(uv)
    device = _pick_device(uv, helem)
    N, E, L = int(mesh.nod2D), int(mesh.elem2D), int(mesh.nl)
    uv_t = _to_device(uv, "uv", device, (E, L, 2))
    he_t = _to_device(helem, "helem", device, (E, L))
    w = torch.empty((N, L), dtype=torch.float64, device=device)
    plan = get_plan(mesh, device)
    with torch.cuda.device(device), plan.on(torch.cuda.current_stream(device)):
        stream = torch.cuda.current_stream(device).cuda_stream
        _lib.check(plan.lib.fct_compute_w(plan.ptr, c_void_p(uv_t.data_ptr()),
                                          c_void_p(he_t.data_ptr()), c_void_p(w.data_ptr()),
                                          c_void_p(stream)), "fct_compute_w")
    return _from_device(w, kind, uv)
