"""ctypes binding of libfesom_fct.so (declared in include/fesom_fct.h).

Loading is explicit and fails loudly: there is no CPU fallback for the FCT path.
``torch`` is imported before the library is opened so that the HIP runtime torch
ships is the one the library binds to (both use the soname libamdhip64.so.7).
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int32, c_int64, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libfesom_fct.so"

FCT_ABI_VERSION = 2
FCT_OK = 0
FCT_E_INVALID = -1
FCT_E_HIP = -2
FCT_E_NOMEM = -3
FCT_E_UNSUPPORTED = -4

FCT_PATH_AUTO = 0
FCT_PATH_PHASED = 1
FCT_PATH_FUSED = 2
FCT_PATH_BIT_PHASED = 1
FCT_PATH_BIT_FUSED = 2

FCT_TUNE_XCD = 1
FCT_TUNE_WAVES_PER_GROUP = 2
FCT_TUNE_VFLUX_EDGES = 3
FCT_TUNE_ADFH_EDGES = 4
FCT_TUNE_XCD_CHUNK = 5
FCT_TUNE_W_EDGE_FLUX = 6
FCT_TUNE_PAIR_TRACERS = 7
FCT_TUNE_ADFH_CHUNK = 8
FCT_TUNE_ADFH_THREADS = 9
FCT_BATCH_SLOTS = 4
FCT_VDIFF_MAX_TRACERS = 8

FCT_PHASE_LO = 1
FCT_PHASE_GRAD = 2
FCT_PHASE_PM = 3
FCT_PHASE_FINAL = 4

EXPORTED_SYMBOLS = ("fct_abi_version", "fct_last_error", "fct_plan_create",
                    "fct_plan_get_info", "fct_plan_destroy", "fct_forward",
                    "fct_forward_ex", "fct_forward_phase", "fct_plan_set_tuning",
                    "fct_batch_create", "fct_batch_set_tuning", "fct_batch_bind",
                    "fct_batch_launch", "fct_batch_graph_launch", "fct_batch_destroy",
                    "fct_backward", "fct_forward_upwind", "fct_impl_vert_diff",
                    "fct_compute_w", "fct_ice_fct_solve", "fct_impl_vert_diff_static", "fct_stage", "fct_tracer_ab",
                    "fct_plan_clone")


class FctMeshDesc(ctypes.Structure):
    _fields_ = [("nod2D", c_int32), ("elem2D", c_int32), ("edge2D", c_int32), ("nl", c_int32),
                ("edges", c_void_p), ("edge_tri", c_void_p), ("edge_cross_dxdy", c_void_p),
                ("edge_dxdy", c_void_p), ("edge_up_dn_tri", c_void_p),
                ("elem_nodes", c_void_p), ("gradient_sca", c_void_p), ("elem_area", c_void_p),
                ("elem_cos", c_void_p), ("nlevels", c_void_p), ("ulevels", c_void_p),
                ("nlevels_nod2D", c_void_p), ("nlevels_nod2D_min", c_void_p),
                ("ulevels_nod2D", c_void_p), ("ulevels_nod2D_max", c_void_p),
                ("nod_in_elem2D_offsets", c_void_p), ("nod_in_elem2D", c_void_p),
                ("area", c_void_p), ("areasvol", c_void_p), ("Z", c_void_p), ("zbar", c_void_p),
                ("max_tracers", c_int32)]


class FctStepArgs(ctypes.Structure):
    _fields_ = [("ntracers", c_int32), ("dt", c_double),
                ("uv", c_void_p), ("w_e", c_void_p), ("helem", c_void_p),
                ("hnode", c_void_p), ("hnode_new", c_void_p),
                ("T", POINTER(c_void_p)), ("T_old", POINTER(c_void_p)),
                ("T_new", POINTER(c_void_p)),
                ("w_i", c_void_p), ("Z3d", c_void_p), ("zbar3", c_void_p)]


FCT_STAGE = {name: i + 1 for i, name in enumerate((
    "TRACER_AB", "EDGE_VFLUX", "ADV_FLUX_HOR", "ADV_FLUX_VER", "FLUX2DTRACER",
    "ALE_RECONSTRUCT", "ADV_FLUX_HOR_HO", "ADV_FLUX_VER_HO", "COMPUTE_FCT_LO",
    "TRACER_GRADIENT_ELEMENTS", "FILL_UP_DN_GRAD", "ZALESAK_LIMIT", "ZALESAK_PM",
    "ZALESAK_APPLY", "FLUX2DTRACER_FCT", "ADV_TRA_VERT_IMPL"))}
FCT_STAGE_MAX_IN = 8


class FctStageArgs(ctypes.Structure):
    _fields_ = [("stage", c_int32), ("dt", c_double), ("in_", c_void_p * FCT_STAGE_MAX_IN),
                ("out", c_void_p * 2)]


class FctAdjointArgs(ctypes.Structure):
    _fields_ = [("g_T_new", POINTER(c_void_p)), ("g_T", POINTER(c_void_p)),
                ("g_T_old", POINTER(c_void_p)), ("g_uv", c_void_p), ("g_w_e", c_void_p),
                ("g_w_i", c_void_p)]


class FctVdiffArgs(ctypes.Structure):
    _fields_ = [("ntracers", c_int32), ("floor_tracer", c_int32), ("dt", c_double),
                ("floor_value", c_double), ("Kv", c_void_p), ("hnode_new", c_void_p),
                ("Z3d", c_void_p), ("T", POINTER(c_void_p)), ("bc_surf", POINTER(c_void_p)),
                ("sw_3d", POINTER(c_void_p)), ("T_new", POINTER(c_void_p))]


class FctIceArgs(ctypes.Structure):
    _fields_ = [("ice_dt", c_double), ("ice_diff", c_double), ("ice_gamma_fct", c_double),
                ("a_ice", c_void_p), ("m_ice", c_void_p), ("m_snow", c_void_p),
                ("u_ice", c_void_p), ("v_ice", c_void_p), ("a_new", c_void_p),
                ("m_new", c_void_p), ("ms_new", c_void_p)]


class FctPhaseBufs(ctypes.Structure):
    _fields_ = [("LO", c_void_p), ("tr_xy", c_void_p), ("fct_plus", c_void_p),
                ("fct_minus", c_void_p)]


class FctPlanInfo(ctypes.Structure):
    _fields_ = [("nod2D", c_int32), ("elem2D", c_int32), ("edge2D", c_int32), ("nl", c_int32),
                ("workspace_bytes", c_int64), ("device", c_int32), ("paths", c_int32),
                ("table_bytes", c_int64), ("table_id", ctypes.c_uint64),
                ("patch_order", c_int32), ("reserved", c_int32)]


_LIB = None


def load(path: Path | None = None, *, require_torch: bool = True):
    """Open the library once and declare the signatures. Raises ``RuntimeError`` if
    the library has not been built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    if require_torch:
        import torch  # noqa: F401  (binds the process to torch's HIP runtime first)
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(str(p))
    lib.fct_abi_version.restype = c_int32
    lib.fct_abi_version.argtypes = []
    lib.fct_last_error.restype = ctypes.c_char_p
    lib.fct_last_error.argtypes = []
    lib.fct_plan_create.restype = c_int32
    lib.fct_plan_create.argtypes = [POINTER(FctMeshDesc), POINTER(c_void_p)]
    lib.fct_plan_get_info.restype = c_int32
    lib.fct_plan_get_info.argtypes = [c_void_p, POINTER(FctPlanInfo)]
    lib.fct_plan_destroy.restype = c_int32
    lib.fct_plan_destroy.argtypes = [c_void_p]
    lib.fct_plan_clone.restype = c_int32
    lib.fct_plan_clone.argtypes = [c_void_p, c_int32, POINTER(c_void_p)]
    lib.fct_forward.restype = c_int32
    lib.fct_forward.argtypes = [c_void_p, POINTER(FctStepArgs), c_void_p]
    lib.fct_forward_ex.restype = c_int32
    lib.fct_forward_ex.argtypes = [c_void_p, POINTER(FctStepArgs), c_int32, c_void_p]
    lib.fct_plan_set_tuning.restype = c_int32
    lib.fct_plan_set_tuning.argtypes = [c_void_p, c_int32, c_int32]
    lib.fct_forward_phase.restype = c_int32
    lib.fct_forward_phase.argtypes = [c_void_p, POINTER(FctStepArgs), c_int32, c_int32,
                                      POINTER(FctPhaseBufs), c_void_p]
    lib.fct_batch_create.restype = c_int32
    lib.fct_batch_create.argtypes = [POINTER(c_void_p), c_int32, POINTER(c_void_p)]
    lib.fct_batch_set_tuning.restype = c_int32
    lib.fct_batch_set_tuning.argtypes = [c_void_p, c_int32, c_int32]
    lib.fct_batch_bind.restype = c_int32
    lib.fct_batch_bind.argtypes = [c_void_p, c_int32, POINTER(FctStepArgs)]
    lib.fct_batch_launch.restype = c_int32
    lib.fct_batch_launch.argtypes = [c_void_p, c_int32, c_void_p]
    lib.fct_batch_graph_launch.restype = c_int32
    lib.fct_batch_graph_launch.argtypes = [c_void_p, c_int32, c_void_p]
    lib.fct_batch_destroy.restype = c_int32
    lib.fct_batch_destroy.argtypes = [c_void_p]
    lib.fct_forward_upwind.restype = c_int32
    lib.fct_forward_upwind.argtypes = [c_void_p, POINTER(FctStepArgs), c_void_p]
    lib.fct_ice_fct_solve.restype = c_int32
    lib.fct_ice_fct_solve.argtypes = [c_void_p, POINTER(FctIceArgs), c_void_p]
    lib.fct_compute_w.restype = c_int32
    lib.fct_compute_w.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
    lib.fct_impl_vert_diff.restype = c_int32
    lib.fct_impl_vert_diff.argtypes = [c_void_p, POINTER(FctVdiffArgs), c_void_p]
    lib.fct_tracer_ab.restype = c_int32
    lib.fct_tracer_ab.argtypes = [c_void_p, c_void_p, c_void_p, ctypes.c_int64, c_void_p]
    lib.fct_stage.restype = c_int32
    lib.fct_stage.argtypes = [c_void_p, POINTER(FctStageArgs), c_void_p]
    lib.fct_impl_vert_diff_static.restype = c_int32
    lib.fct_impl_vert_diff_static.argtypes = [c_void_p, POINTER(FctVdiffArgs), c_void_p]
    lib.fct_backward.restype = c_int32
    lib.fct_backward.argtypes = [c_void_p, POINTER(FctStepArgs), POINTER(FctAdjointArgs),
                                 c_void_p]
    ver = lib.fct_abi_version()
    if ver != FCT_ABI_VERSION:
        raise RuntimeError(f"libfesom_fct ABI {ver} != expected {FCT_ABI_VERSION}")
    if path is None:
        _LIB = lib
    return lib


MESH_INT_FIELDS = ("edges", "edge_tri", "edge_up_dn_tri", "elem_nodes", "nlevels", "ulevels",
                   "nlevels_nod2D", "nlevels_nod2D_min", "ulevels_nod2D", "ulevels_nod2D_max",
                   "nod_in_elem2D_offsets", "nod_in_elem2D")
MESH_F64_FIELDS = ("edge_cross_dxdy", "edge_dxdy", "gradient_sca", "elem_area", "elem_cos",
                   "area", "areasvol", "Z", "zbar")


def mesh_desc_from(mesh, max_tracers: int = 4):
    """``fct_mesh_desc`` over contiguous host copies of the mesh fields the FCT reads
    (the reference's Mesh names, fesom_jax/mesh.py:62-129). Returns the descriptor and
    the arrays it points into; keep both alive until ``fct_plan_create`` returns."""
    import numpy as np
    N, E, D, L = int(mesh.nod2D), int(mesh.elem2D), int(mesh.edge2D), int(mesh.nl)
    host = {}
    for k in MESH_INT_FIELDS:
        host[k] = np.ascontiguousarray(np.asarray(getattr(mesh, k)), dtype=np.int32)
    for k in MESH_F64_FIELDS:
        host[k] = np.ascontiguousarray(np.asarray(getattr(mesh, k)), dtype=np.float64)
    # the library reads these element counts from the pointers: refuse any mismatch
    offsets = host["nod_in_elem2D_offsets"]
    want = {"edges": (D, 2), "edge_tri": (D, 2), "edge_up_dn_tri": (D, 2), "elem_nodes": (E, 3),
            "nlevels": (E,), "ulevels": (E,), "nlevels_nod2D": (N,), "nlevels_nod2D_min": (N,),
            "ulevels_nod2D": (N,), "ulevels_nod2D_max": (N,), "nod_in_elem2D_offsets": (N + 1,),
            "edge_cross_dxdy": (D, 4), "edge_dxdy": (D, 2), "gradient_sca": (E, 6),
            "elem_area": (E,), "elem_cos": (E,), "area": (N, L), "areasvol": (N, L),
            "Z": (L - 1,), "zbar": (L,)}
    for k, shape in want.items():
        if host[k].shape != shape:
            raise ValueError(f"mesh field {k} has shape {host[k].shape}, expected {shape} "
                             f"(nod2D={N}, elem2D={E}, edge2D={D}, nl={L})")
    if offsets[0] != 0 or np.any(np.diff(offsets) < 0) or host["nod_in_elem2D"].shape != (
            int(offsets[-1]),):
        raise ValueError("nod_in_elem2D_offsets must rise from 0 to len(nod_in_elem2D)")
    desc = FctMeshDesc(nod2D=int(mesh.nod2D), elem2D=int(mesh.elem2D),
                       edge2D=int(mesh.edge2D), nl=int(mesh.nl), max_tracers=int(max_tracers))
    for k, a in host.items():
        setattr(desc, k, a.ctypes.data)
    return desc, host


def check(rc: int, what: str) -> None:
    """Map a C status to the Python exception the reference's conventions imply."""
    if rc == FCT_OK:
        return
    msg = (_LIB.fct_last_error() or b"").decode("utf-8", "replace") if _LIB else ""
    text = f"{what}: {msg}"
    if rc == FCT_E_INVALID:
        raise ValueError(text)
    if rc == FCT_E_UNSUPPORTED:
        raise NotImplementedError(text)
    if rc == FCT_E_NOMEM:
        raise MemoryError(text)
    raise RuntimeError(text)
