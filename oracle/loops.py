"""ctypes front end of the C loop oracle (oracle/fct_loops.c), test infrastructure only:
one FCT tracer step with the original C kernel's loop semantics, as restated by the
reference's loop oracle (fesom_jax/tests/test_tracers.py:325-516)."""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int

import numpy as np

from . import build as _build

_INT_FIELDS = ("edges", "edge_tri", "edge_up_dn_tri", "elem_nodes", "ulevels", "nlevels",
               "ulevels_nod2D", "nlevels_nod2D", "ulevels_nod2D_max", "nlevels_nod2D_min",
               "nod_in_elem2D_offsets", "nod_in_elem2D")
_F64_FIELDS = ("edge_cross_dxdy", "edge_dxdy", "gradient_sca", "elem_area", "elem_cos",
               "area", "areasvol", "Z", "zbar")


class _LoopMesh(ctypes.Structure):
    _fields_ = ([("nod2D", c_int), ("elem2D", c_int), ("edge2D", c_int), ("nl", c_int)]
                + [(f, POINTER(c_int)) for f in _INT_FIELDS]
                + [(f, POINTER(c_double)) for f in _F64_FIELDS])


class _IceCfg(ctypes.Structure):
    _fields_ = [("dt", c_double), ("diff", c_double), ("gamma", c_double)]


_LIBS = {}


def _lib(openmp: bool = False):
    if openmp not in _LIBS:
        _LIBS[openmp] = lib = ctypes.CDLL(str(_build.ensure(openmp)))
        _LIB = lib
        _LIB.fct_loops_step.restype = c_int
        _LIB.fct_loops_step.argtypes = [POINTER(_LoopMesh)] + [POINTER(c_double)] * 7 + [
            c_double, POINTER(c_double)]
        _LIB.loops_ice_fct.restype = c_int
        _LIB.loops_ice_fct.argtypes = [POINTER(_LoopMesh), POINTER(_IceCfg)] + [
            POINTER(c_double)] * 8
        _LIB.loops_compute_w.restype = c_int
        _LIB.loops_compute_w.argtypes = [POINTER(_LoopMesh)] + [POINTER(c_double)] * 3
    return _LIBS[openmp]


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _mesh_struct(mesh):
    keep = {}
    kw = {}
    for f in _INT_FIELDS:
        keep[f] = np.ascontiguousarray(getattr(mesh, f), dtype=np.int32)
        kw[f] = keep[f].ctypes.data_as(POINTER(c_int))
    for f in _F64_FIELDS:
        keep[f] = _f64(getattr(mesh, f))
        kw[f] = keep[f].ctypes.data_as(POINTER(c_double))
    m = _LoopMesh(nod2D=int(mesh.nod2D), elem2D=int(mesh.elem2D), edge2D=int(mesh.edge2D),
                  nl=int(mesh.nl), **kw)
    return m, keep


def advect_one_fct(mesh, uv, w_e, helem, hnode, hnode_new, T, T_old, *, dt, openmp=False):
    """T_new of one FCT step (same arguments as the reference's advect_one_fct);
    ``openmp`` runs the OpenMP build (threads from OMP_NUM_THREADS, same bits)."""
    m, keep = _mesh_struct(mesh)
    args = [_f64(x) for x in (uv, w_e, helem, hnode, hnode_new, T, T_old)]
    out = np.empty((int(mesh.nod2D), int(mesh.nl)))
    P = lambda a: a.ctypes.data_as(POINTER(c_double))
    rc = _lib(openmp).fct_loops_step(ctypes.byref(m), *[P(a) for a in args], float(dt), P(out))
    if rc != 0:
        raise MemoryError("fct_loops_step: allocation failed")
    return out


def compute_w(mesh, uv, helem):
    """w [nod2D, nl] with the original C kernel's loops (fesom_jax/ale.py:309-365)."""
    m, keep = _mesh_struct(mesh)
    uv, helem = _f64(uv), _f64(helem)
    out = np.empty((int(mesh.nod2D), int(mesh.nl)))
    P = lambda a: a.ctypes.data_as(POINTER(c_double))
    _lib().loops_compute_w(ctypes.byref(m), P(uv), P(helem), P(out))
    return out


def ice_fct_solve(cfg, mesh, a_ice, m_ice, m_snow, u_ice, v_ice):
    """(a_new, m_new, ms_new) of one sea-ice FCT step with element loops
    (fesom_jax/ice_adv.py:177-194)."""
    m, keep = _mesh_struct(mesh)
    c = _IceCfg(dt=float(cfg.ice_dt), diff=float(cfg.ice_diff), gamma=float(cfg.ice_gamma_fct))
    ins = [_f64(x) for x in (a_ice, m_ice, m_snow, u_ice, v_ice)]
    outs = [np.empty(int(mesh.nod2D)) for _ in range(3)]
    P = lambda a: a.ctypes.data_as(POINTER(c_double))
    rc = _lib().loops_ice_fct(ctypes.byref(m), ctypes.byref(c), *[P(a) for a in ins],
                              *[P(o) for o in outs])
    if rc != 0:
        raise MemoryError("loops_ice_fct: allocation failed")
    return tuple(outs)
