"""Drop-in FCT tracer advection on the GPU (the reference's
fesom_jax/tracer_adv.py:532-578 ``advect_one_fct``).

``advect_one_fct`` has the reference's name, arguments, defaults and return value:
``(T_new, T_old_new)`` with ``T_old_new`` the very ``T`` passed in
(fesom_jax/tracer_adv.py:59-66, 549). It accepts torch tensors on a ROCm device,
objects exporting DLPack (e.g. JAX arrays on the GPU, exchanged zero-copy) and numpy
arrays (copied to the current device and the result copied back). Every field must be
float64 with the reference's shapes: node fields ``[nod2D, nl]``, ``uv``
``[elem2D, nl, 2]``, ``helem`` ``[elem2D, nl]``.

The work runs on torch's current HIP stream of the tensors' device, through the C ABI
of libfesom_fct.so. ``exch`` (the reference's halo-exchange hook) is honoured: it is
called on the low-order solution, the element gradients and the limiter sums, in the
reference's order (tracer_adv.py:559, 562, 438-439), between phases of the kernel
pipeline; ``exch=None`` runs the single-call path.

There is no CPU fallback: without a ROCm GPU or the built library the call raises.
"""

from __future__ import annotations

import contextlib
import ctypes
import weakref
from ctypes import c_void_p

import numpy as np

from . import _lib
from .config import DT_DEFAULT

__all__ = ["advect_one_fct", "advect_fct", "FctPlan", "get_plan", "DT_DEFAULT",
           "tracer_ab", "edge_vflux", "adv_flux_hor", "adv_flux_ver", "flux2dtracer",
           "ale_reconstruct", "adv_flux_hor_ho", "adv_flux_ver_ho", "compute_fct_lo",
           "tracer_gradient_elements", "fill_up_dn_grad", "zalesak_limit",
           "flux2dtracer_fct", "adv_tra_vert_impl"]


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("fesom_jax_amd.advect_one_fct needs a ROCm GPU; none is visible")
    return torch


# --------------------------------------------------------------------------
# plan (device topology) cache
# --------------------------------------------------------------------------


class FctPlan:
    """Device-resident topology + workspace for one mesh on one device."""

    def __init__(self, mesh, device=None, max_tracers: int = 4, *, share=None):
        """``share``: a live plan of the same mesh and device whose device tables the
        new plan shares (fct_plan_clone); the new plan then only allocates workspace."""
        torch = _torch()
        self.lib = _lib.load()
        dev = torch.device("cuda", torch.cuda.current_device() if device is None
                           else torch.device(device).index or 0)
        self.device = dev
        self.nod2D = int(mesh.nod2D)
        self.elem2D = int(mesh.elem2D)
        self.edge2D = int(mesh.edge2D)
        self.nl = int(mesh.nl)
        ptr = c_void_p()
        if share is not None:
            if (share.device != dev or share.nod2D != self.nod2D or share.nl != self.nl
                    or share.edge2D != self.edge2D or share.elem2D != self.elem2D):
                raise ValueError("share must be a plan of the same mesh and device")
            with torch.cuda.device(dev):
                _lib.check(self.lib.fct_plan_clone(share.ptr, int(max_tracers),
                                                   ctypes.byref(ptr)), "fct_plan_clone")
        else:
            desc, _host = _lib.mesh_desc_from(mesh, max_tracers)
            with torch.cuda.device(dev):
                _lib.check(self.lib.fct_plan_create(ctypes.byref(desc), ctypes.byref(ptr)),
                           "fct_plan_create")
        self.ptr = ptr
        self._finalizer = weakref.finalize(self, self.lib.fct_plan_destroy, ptr)
        info = _lib.FctPlanInfo()
        _lib.check(self.lib.fct_plan_get_info(ptr, ctypes.byref(info)), "fct_plan_get_info")
        self.workspace_bytes = int(info.workspace_bytes)
        self.fused_available = bool(info.paths & _lib.FCT_PATH_BIT_FUSED)
        self._last = None          # stream of the last launch on this plan
        self._event = None         # recorded after that launch

    @contextlib.contextmanager
    def on(self, stream):
        """Order a launch on ``stream`` after the plan's previous launches: calls on one
        plan share its workspace, so when the stream changes the new stream first waits
        for the last one's work; the launch is then recorded for the next caller."""
        torch = _torch()
        if self._last is not None and self._last != stream.cuda_stream:
            stream.wait_event(self._event)
        yield
        if self._event is None:
            self._event = torch.cuda.Event()
        self._event.record(stream)
        self._last = stream.cuda_stream

    def _info(self):
        info = _lib.FctPlanInfo()
        _lib.check(self.lib.fct_plan_get_info(self.ptr, ctypes.byref(info)), "fct_plan_get_info")
        return info

    def device_bytes(self) -> int:
        """Device memory the plan keeps alive now (its mesh tables, shared with the
        plans cloned from it, and its own workspace, including what the phased,
        adjoint, diffusion and sea-ice entry points allocated on first use)."""
        return int(self._info().workspace_bytes)

    def own_bytes(self) -> int:
        """The plan's own workspace (device_bytes without the shared mesh tables)."""
        info = self._info()
        return int(info.workspace_bytes - info.table_bytes)

    def tables(self):
        """``(table_id, table_bytes)``: plans with the same id share their tables."""
        info = self._info()
        return int(info.table_id), int(info.table_bytes)

    @property
    def patch_order(self) -> bool:
        """Whether the mesh's numbering was found scattered, so that the node and
        vertical-flux passes visit it in the edge patches' breadth-first order."""
        return bool(self._info().patch_order)

    def close(self):
        self._finalizer()


_PLANS: dict = {}


def get_plan(mesh, device) -> FctPlan:
    """The cached plan of ``mesh`` on ``device`` (keyed by object identity; dropped
    when the mesh object is garbage collected)."""
    key = (id(mesh), str(device))
    plan = _PLANS.get(key)
    if plan is None:
        plan = FctPlan(mesh, device)
        _PLANS[key] = plan
        try:
            weakref.finalize(mesh, _PLANS.pop, key, None)
        except TypeError:
            pass
    return plan


# --------------------------------------------------------------------------
# array interchange
# --------------------------------------------------------------------------
class _Kind:
    TORCH, NUMPY, DLPACK = "torch", "numpy", "dlpack"


def _kind_of(x):
    torch = _torch()
    if isinstance(x, torch.Tensor):
        return _Kind.TORCH
    if isinstance(x, np.ndarray):
        return _Kind.NUMPY
    if hasattr(x, "__dlpack__"):
        return _Kind.DLPACK
    return _Kind.NUMPY


def _to_device(x, name, device, shape):
    torch = _torch()
    if isinstance(x, torch.Tensor):
        t = x
    elif isinstance(x, np.ndarray) or not hasattr(x, "__dlpack__"):
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(x), dtype=np.float64))
    else:
        t = torch.from_dlpack(x)
    if t.dtype != torch.float64:
        raise ValueError(f"{name}: dtype {t.dtype} is not float64 (the FCT path is FP64)")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)} != expected {tuple(shape)}")
    if t.device != device:
        if t.device.type == "cpu":
            t = t.to(device)
        else:
            raise ValueError(f"{name}: on {t.device}, expected {device}")
    return t.contiguous()


def _from_device(t, kind, like):
    if kind == _Kind.TORCH:
        return t
    if kind == _Kind.NUMPY:
        return t.cpu().numpy()
    mod = type(like).__module__.split(".")[0]
    if mod in ("jax", "jaxlib"):
        import jax.dlpack  # type: ignore
        return jax.dlpack.from_dlpack(t)
    return t


def _pick_device(*xs):
    torch = _torch()
    for x in xs:
        if isinstance(x, torch.Tensor) and x.is_cuda:
            return x.device
    return torch.device("cuda", torch.cuda.current_device())


def _ptr_array(ts):
    arr = (c_void_p * len(ts))()
    for i, t in enumerate(ts):
        arr[i] = t.data_ptr()
    return arr


def _run(mesh, uv, w_e, helem, hnode, hnode_new, Ts, T_olds, dt, exch, Z3d, zbar3, w_i,
         path=_lib.FCT_PATH_AUTO, plan=None):
    torch = _torch()
    device = _pick_device(T_olds[0], Ts[0], uv, w_e)
    N, E, L = int(mesh.nod2D), int(mesh.elem2D), int(mesh.nl)
    nshape, eshape = (N, L), (E, L)
    uv_t = _to_device(uv, "uv", device, (E, L, 2))
    w_t = _to_device(w_e, "w_e", device, nshape)
    he_t = _to_device(helem, "helem", device, eshape)
    hn_t = _to_device(hnode, "hnode", device, nshape)
    hnn_t = _to_device(hnode_new, "hnode_new", device, nshape)
    T_t = [_to_device(T, f"T[{j}]", device, nshape) for j, T in enumerate(Ts)]
    To_t = [_to_device(T, f"T_old[{j}]", device, nshape) for j, T in enumerate(T_olds)]
    wi_t = None if w_i is None else _to_device(w_i, "w_i", device, nshape)
    if (Z3d is None) != (zbar3 is None):
        raise ValueError("Z3d and zbar3 must be given together")
    z3_t = None if Z3d is None else _to_device(Z3d, "Z3d", device, nshape)
    zb3_t = None if zbar3 is None else _to_device(zbar3, "zbar3", device, nshape)
    out = [torch.empty(nshape, dtype=torch.float64, device=device) for _ in Ts]

    if plan is None:
        plan = get_plan(mesh, device)
    elif plan.device != device or plan.nod2D != N or plan.nl != L:
        raise ValueError("plan does not belong to this mesh and device")
    Tp, Top, Tnp = _ptr_array(T_t), _ptr_array(To_t), _ptr_array(out)
    args = _lib.FctStepArgs(
        ntracers=len(Ts), dt=float(dt), uv=uv_t.data_ptr(), w_e=w_t.data_ptr(),
        helem=he_t.data_ptr(), hnode=hn_t.data_ptr(), hnode_new=hnn_t.data_ptr(),
        T=ctypes.cast(Tp, ctypes.POINTER(c_void_p)),
        T_old=ctypes.cast(Top, ctypes.POINTER(c_void_p)),
        T_new=ctypes.cast(Tnp, ctypes.POINTER(c_void_p)),
        w_i=None if wi_t is None else wi_t.data_ptr(),
        Z3d=None if z3_t is None else z3_t.data_ptr(),
        zbar3=None if zb3_t is None else zb3_t.data_ptr())
    with torch.cuda.device(device), plan.on(torch.cuda.current_stream(device)):
        stream = torch.cuda.current_stream(device).cuda_stream
        if path == _UPWIND:
            _lib.check(plan.lib.fct_forward_upwind(plan.ptr, ctypes.byref(args), c_void_p(stream)),
                       "fct_forward_upwind")
        elif exch is None:
            _lib.check(plan.lib.fct_forward_ex(plan.ptr, ctypes.byref(args), int(path),
                                               c_void_p(stream)), "fct_forward")
        else:
            for j in range(len(Ts)):
                _run_phased(plan, args, j, exch, device, N, E, L, stream)
    return out


def _exch_field(exch, t, kind, shape, device, name):
    """Apply the caller's exch to an intermediate and bring the result back to a
    contiguous float64 device tensor of the same shape."""
    r = exch(t, kind)
    if r is t:
        return t
    return _to_device(r, name, device, shape)


def _run_phased(plan, args, j, exch, device, N, E, L, stream):
    torch = _torch()
    f64 = dict(dtype=torch.float64, device=device)
    LO = torch.empty((N, L), **f64)
    tr = torch.empty((E, L, 2), **f64)
    fp = torch.empty((N, L), **f64)
    fm = torch.empty((N, L), **f64)
    s = c_void_p(stream)

    def bufs():
        return _lib.FctPhaseBufs(LO=LO.data_ptr(), tr_xy=tr.data_ptr(),
                                 fct_plus=fp.data_ptr(), fct_minus=fm.data_ptr())

    b = bufs()
    _lib.check(plan.lib.fct_forward_phase(plan.ptr, ctypes.byref(args), j, _lib.FCT_PHASE_LO,
                                          ctypes.byref(b), s), "phase LO")
    LO = _exch_field(exch, LO, "nod", (N, L), device, "exch(LO)")
    b = bufs()
    _lib.check(plan.lib.fct_forward_phase(plan.ptr, ctypes.byref(args), j,
                                          _lib.FCT_PHASE_GRAD, ctypes.byref(b), s), "phase GRAD")
    tr = _exch_field(exch, tr, "elem", (E, L, 2), device, "exch(tr_xy)")
    b = bufs()
    _lib.check(plan.lib.fct_forward_phase(plan.ptr, ctypes.byref(args), j, _lib.FCT_PHASE_PM,
                                          ctypes.byref(b), s), "phase PM")
    fp = _exch_field(exch, fp, "nod", (N, L), device, "exch(fct_plus)")
    fm = _exch_field(exch, fm, "nod", (N, L), device, "exch(fct_minus)")
    b = bufs()
    _lib.check(plan.lib.fct_forward_phase(plan.ptr, ctypes.byref(args), j,
                                          _lib.FCT_PHASE_FINAL, ctypes.byref(b), s),
               "phase FINAL")


_UPWIND = -1   # _run path code of the upwind-only step


# --------------------------------------------------------------------------
# public entry points
# --------------------------------------------------------------------------
def advect_one(mesh, uv, w_e, helem, hnode, hnode_new, T, T_old, *, dt: float = DT_DEFAULT):
    """One tracer's first-order upwind advection of the AB2 field + ALE reconstruction;
    returns ``(T_new, T)`` (fesom_jax/tracer_adv.py:154-165)."""
    kind = _kind_of(T)
    out = _run(mesh, uv, w_e, helem, hnode, hnode_new, [T], [T_old], dt, None, None, None,
               None, _UPWIND)
    return _from_device(out[0], kind, T), T


def advect_one_fct(mesh, uv, w_e, helem, hnode, hnode_new, T, T_old, *,
                   dt: float = DT_DEFAULT, exch=None, Z3d=None, zbar3=None, w_i=None):
    """One tracer's FCT advection + ALE reconstruction; returns ``(T_new, T)``
    (fesom_jax/tracer_adv.py:532-578)."""
    kind = _kind_of(T)
    out = _run(mesh, uv, w_e, helem, hnode, hnode_new, [T], [T_old], dt, exch, Z3d, zbar3,
               w_i)
    return _from_device(out[0], kind, T), T


def advect_fct(mesh, uv, w_e, helem, hnode, hnode_new, Ts, T_olds, *,
               dt: float = DT_DEFAULT, exch=None, Z3d=None, zbar3=None, w_i=None,
               path: str = "auto", plan=None):
    """K tracers sharing one flow in one call (the two calls of fesom_jax/step.py:409-416
    fused): returns ``[(T_new_k, T_k)]``. ``path`` selects the kernel path ("auto",
    "fused", "phased"; all give the same bits); ``plan`` an ``FctPlan`` of the mesh to
    run on instead of the mesh's cached one. Extension with no reference
    counterpart."""
    if len(Ts) != len(T_olds) or not Ts:
        raise ValueError("Ts and T_olds must be non-empty and of equal length")
    codes = {"auto": _lib.FCT_PATH_AUTO, "fused": _lib.FCT_PATH_FUSED,
             "phased": _lib.FCT_PATH_PHASED}
    if path not in codes:
        raise ValueError(f"path must be one of {sorted(codes)}")
    out = _run(mesh, uv, w_e, helem, hnode, hnode_new, list(Ts), list(T_olds), dt, exch, Z3d,
               zbar3, w_i, codes[path], plan=plan)
    return [(_from_device(o, _kind_of(T), T), T) for o, T in zip(out, Ts)]


# --------------------------------------------------------------------------
# the reference's stage functions, one GPU stage each (fct_stage in the C ABI)
# --------------------------------------------------------------------------
def _stage(mesh, name, ins, out_shapes, *, dt=0.0):
    """Run stage ``name`` on ``ins`` (a list of (array or None, name, shape)); returns
    the outputs in the kind of the first input."""
    torch = _torch()
    given = [x for x, _, _ in ins if x is not None]
    device = _pick_device(*given)
    ts = [None if x is None else _to_device(x, nm, device, shape) for x, nm, shape in ins]
    outs = [torch.empty(shape, dtype=torch.float64, device=device) for shape in out_shapes]
    plan = get_plan(mesh, device)
    a = _lib.FctStageArgs(stage=_lib.FCT_STAGE[name], dt=float(dt))
    for i, t in enumerate(ts):
        a.in_[i] = None if t is None else t.data_ptr()
    for i, o in enumerate(outs):
        a.out[i] = o.data_ptr()
    with torch.cuda.device(device), plan.on(torch.cuda.current_stream(device)):
        stream = torch.cuda.current_stream(device).cuda_stream
        _lib.check(plan.lib.fct_stage(plan.ptr, ctypes.byref(a), c_void_p(stream)),
                   f"fct_stage({name})")
    kind, like = _kind_of(ins[0][0]), ins[0][0]
    return [_from_device(o, kind, like) for o in outs]


def _shapes(mesh):
    N, E, D, L = int(mesh.nod2D), int(mesh.elem2D), int(mesh.edge2D), int(mesh.nl)
    return (N, L), (E, L), (D, L), N, E, D, L


def tracer_ab(values, valuesold):
    """``(ttfAB, T_old_new)``: ``ttfAB = -(0.5+0.1)*valuesold + (1.5+0.1)*values`` on the
    GPU, ``T_old_new`` the input ``values`` itself (fesom_jax/tracer_adv.py:59-66)."""
    torch = _torch()
    device = _pick_device(values, valuesold)
    shape = tuple(np.shape(values)) if not isinstance(values, torch.Tensor) else tuple(values.shape)
    v = _to_device(values, "values", device, shape)
    vo = _to_device(valuesold, "valuesold", device, shape)
    out = torch.empty(shape, dtype=torch.float64, device=device)
    lib = _lib.load()
    with torch.cuda.device(device):
        stream = torch.cuda.current_stream(device).cuda_stream
        _lib.check(lib.fct_tracer_ab(c_void_p(v.data_ptr()), c_void_p(vo.data_ptr()),
                                     c_void_p(out.data_ptr()), int(out.numel()),
                                     c_void_p(stream)), "fct_tracer_ab")
    return _from_device(out, _kind_of(values), values), values


def edge_vflux(mesh, uv, helem):
    """Volume flux through each edge-level (fesom_jax/tracer_adv.py:69-93)."""
    ns, es, ds, N, E, D, L = _shapes(mesh)
    return _stage(mesh, "EDGE_VFLUX", [(uv, "uv", (E, L, 2)), (helem, "helem", es)], [ds])[0]


def adv_flux_hor(mesh, uv, helem, ttf):
    """Low-order upwind horizontal flux (fesom_jax/tracer_adv.py:96-105)."""
    ns, es, ds, N, E, D, L = _shapes(mesh)
    return _stage(mesh, "ADV_FLUX_HOR", [(uv, "uv", (E, L, 2)), (helem, "helem", es),
                                         (ttf, "ttf", ns)], [ds])[0]


def adv_flux_ver(mesh, w_e, ttf):
    """Low-order upwind vertical flux at the interfaces (fesom_jax/tracer_adv.py:108-120)."""
    ns = _shapes(mesh)[0]
    return _stage(mesh, "ADV_FLUX_VER", [(w_e, "w_e", ns), (ttf, "ttf", ns)], [ns])[0]


def flux2dtracer(mesh, flux_h, flux_v, *, dt: float = DT_DEFAULT):
    """Tracer change of the upwind driver (fesom_jax/tracer_adv.py:123-139)."""
    ns, es, ds = _shapes(mesh)[:3]
    return _stage(mesh, "FLUX2DTRACER", [(flux_h, "flux_h", ds), (flux_v, "flux_v", ns)], [ns],
                  dt=dt)[0]


def ale_reconstruct(mesh, T, del_ttf, hnode, hnode_new):
    """``T + del_ttf / hnode_new`` on the wet layers with the thickness change
    (fesom_jax/tracer_adv.py:142-151)."""
    ns = _shapes(mesh)[0]
    return _stage(mesh, "ALE_RECONSTRUCT", [(T, "T", ns), (del_ttf, "del_ttf", ns),
                                            (hnode, "hnode", ns), (hnode_new, "hnode_new", ns)],
                  [ns])[0]


def adv_flux_hor_ho(mesh, uv, helem, ttf, eud):
    """MFCT high-order horizontal flux from the gradient slots ``eud``
    (fesom_jax/tracer_adv.py:196-225)."""
    ns, es, ds, N, E, D, L = _shapes(mesh)
    return _stage(mesh, "ADV_FLUX_HOR_HO", [(uv, "uv", (E, L, 2)), (helem, "helem", es),
                                            (ttf, "ttf", ns), (eud, "eud", (D, L, 4))], [ds])[0]


def adv_flux_ver_ho(mesh, w_e, ttf, Z3d=None, zbar3=None):
    """QR4C high-order vertical flux (fesom_jax/tracer_adv.py:241-295)."""
    ns = _shapes(mesh)[0]
    if (Z3d is None) != (zbar3 is None):
        raise ValueError("Z3d and zbar3 must be given together")
    return _stage(mesh, "ADV_FLUX_VER_HO", [(w_e, "w_e", ns), (ttf, "ttf", ns),
                                            (Z3d, "Z3d", ns), (zbar3, "zbar3", ns)], [ns])[0]


def compute_fct_lo(mesh, flux_h, flux_v, T, hnode, hnode_new, *, dt: float = DT_DEFAULT):
    """Low-order solution (fesom_jax/tracer_adv.py:298-313)."""
    ns, es, ds = _shapes(mesh)[:3]
    return _stage(mesh, "COMPUTE_FCT_LO", [(flux_h, "flux_h", ds), (flux_v, "flux_v", ns),
                                           (T, "T", ns), (hnode, "hnode", ns),
                                           (hnode_new, "hnode_new", ns)], [ns], dt=dt)[0]


def tracer_gradient_elements(mesh, ttf):
    """Per element-level gradient ``[elem2D, nl, 2]`` (fesom_jax/tracer_adv.py:316-326)."""
    ns, es, ds, N, E, D, L = _shapes(mesh)
    return _stage(mesh, "TRACER_GRADIENT_ELEMENTS", [(ttf, "ttf", ns)], [(E, L, 2)])[0]


def fill_up_dn_grad(mesh, tr_xy):
    """The four MFCT gradient slots ``[edge2D, nl, 4]`` (fesom_jax/tracer_adv.py:344-368)."""
    ns, es, ds, N, E, D, L = _shapes(mesh)
    return _stage(mesh, "FILL_UP_DN_GRAD", [(tr_xy, "tr_xy", (E, L, 2))], [(D, L, 4)])[0]


def zalesak_limit(mesh, T, LO, adf_h, adf_v, hnode_new, *, dt: float = DT_DEFAULT, exch=None):
    """Zalesak limiter: ``(adf_h_lim, adf_v_lim)`` (fesom_jax/tracer_adv.py:371-466).
    With ``exch`` the P+/P- sums go through it (``"nod"``, :438-439) before the factors."""
    ns, es, ds = _shapes(mesh)[:3]
    ins = [(T, "T", ns), (LO, "LO", ns), (adf_h, "adf_h", ds), (adf_v, "adf_v", ns),
           (hnode_new, "hnode_new", ns)]
    if exch is None:
        return tuple(_stage(mesh, "ZALESAK_LIMIT", ins, [ds, ns], dt=dt))
    fp, fm = _stage(mesh, "ZALESAK_PM", ins, [ns, ns], dt=dt)
    fp, fm = exch(fp, "nod"), exch(fm, "nod")
    return tuple(_stage(mesh, "ZALESAK_APPLY", ins + [(fp, "fct_plus", ns),
                                                      (fm, "fct_minus", ns)], [ds, ns], dt=dt))


def flux2dtracer_fct(mesh, T, LO, adf_h, adf_v, hnode, hnode_new, *, dt: float = DT_DEFAULT):
    """``del_ttf`` of the FCT path from the limited fluxes (fesom_jax/tracer_adv.py:469-484)."""
    ns, es, ds = _shapes(mesh)[:3]
    return _stage(mesh, "FLUX2DTRACER_FCT", [(T, "T", ns), (LO, "LO", ns), (adf_h, "adf_h", ds),
                                             (adf_v, "adf_v", ns), (hnode, "hnode", ns),
                                             (hnode_new, "hnode_new", ns)], [ns], dt=dt)[0]


def adv_tra_vert_impl(mesh, w_i, T, hnode_new, *, dt: float = DT_DEFAULT):
    """Implicit vertical advection by the split velocity ``w_i``
    (fesom_jax/tracer_adv.py:487-529)."""
    ns = _shapes(mesh)[0]
    return _stage(mesh, "ADV_TRA_VERT_IMPL", [(w_i, "w_i", ns), (T, "T", ns),
                                              (hnode_new, "hnode_new", ns)], [ns], dt=dt)[0]
