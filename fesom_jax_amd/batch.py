"""Many independent FCT instances stepped together.

The reference advances one mesh per call (``fesom_jax/tracer_adv.py:532-578``); there
is no batched counterpart. Here each instance keeps its mesh plan (with its own
workspace), flow fields and tracers resident on the device, and a ``FctBatch`` runs one
FCT step of every instance with one launch per pass over the concatenated task ranges
(``fct_batch_*`` in ``include/fesom_fct.h``), optionally replayed as a HIP graph.
Results are bit-identical to stepping each instance alone with ``fct_forward``.

State rotation: a step maps ``(T_old, T, spare) -> (T, T_new, ...)``; with three
device buffers per tracer the buffer roles repeat every three steps, so the batch binds
three argument slots once and then only launches.
"""

from __future__ import annotations

import ctypes
from ctypes import c_void_p
from dataclasses import dataclass, field

from . import _lib
from .tracer_adv import FctPlan, _ptr_array, _torch


@dataclass
class Instance:
    """One FCT problem resident on the device. ``bufs[j]`` holds three float64
    ``[N, L]`` buffers for tracer j; ``phase`` says which of them is T_old, T and the
    spare the next step writes (T_old = bufs[j][phase], T = bufs[j][phase+1 mod 3])."""

    plan: FctPlan
    uv: object
    w_e: object
    helem: object
    hnode: object
    hnode_new: object
    bufs: list
    dt: float
    updates: int = 0          # node-level updates per step (K * sum_Ln)
    phase: int = 0
    _keep: list = field(default_factory=list, repr=False)

    @classmethod
    def from_case(cls, mesh, case, device, *, max_tracers=None, plan=None):
        """Upload a case dict (``uv, w_e, helem, hnode, hnode_new, T, T_old, dt``; numpy
        or torch) and build a private plan for it (or take ``plan``, which must be
        built for this mesh and not be used by another live instance)."""
        torch = _torch()
        t = lambda a: (a if isinstance(a, torch.Tensor) else torch.from_numpy(a)).to(
            device=device, dtype=torch.float64).contiguous()
        K = len(case["T"])
        N, E, L = int(mesh.nod2D), int(mesh.elem2D), int(mesh.nl)
        want = {"uv": (E, L, 2), "w_e": (N, L), "helem": (E, L), "hnode": (N, L),
                "hnode_new": (N, L)}
        for k, shape in want.items():
            if tuple(case[k].shape) != shape:
                raise ValueError(f"{k} has shape {tuple(case[k].shape)}, expected {shape}")
        if len(case["T_old"]) != K or K < 1:
            raise ValueError("T and T_old must hold the same number (>= 1) of tracers")
        for j, (T, To) in enumerate(zip(case["T"], case["T_old"])):
            if tuple(T.shape) != (N, L) or tuple(To.shape) != (N, L):
                raise ValueError(f"tracer {j} has shape {tuple(T.shape)} / {tuple(To.shape)}, "
                                 f"expected {(N, L)}")
        if plan is None:
            plan = FctPlan(mesh, device, max_tracers=max_tracers or K)
        elif plan.nod2D != N or plan.nl != L or plan.elem2D != E:
            raise ValueError("plan was built for another mesh")
        bufs = []
        for T, To in zip(case["T"], case["T_old"]):
            Tt, Tot = t(T), t(To)
            bufs.append([Tot.clone(), Tt.clone(), torch.empty_like(Tt)])
        from .perfmodel import updates_per_step
        return cls(plan=plan, uv=t(case["uv"]), w_e=t(case["w_e"]), helem=t(case["helem"]),
                   hnode=t(case["hnode"]), hnode_new=t(case["hnode_new"]), bufs=bufs,
                   dt=float(case["dt"]), updates=updates_per_step(mesh, K))

    @property
    def ntracers(self) -> int:
        return len(self.bufs)

    def roles(self, phase=None):
        """(T_old, T, T_new) tensors of every tracer at rotation phase ``phase``."""
        p = self.phase if phase is None else phase
        return ([b[p % 3] for b in self.bufs], [b[(p + 1) % 3] for b in self.bufs],
                [b[(p + 2) % 3] for b in self.bufs])

    def T(self):
        """Current tracer fields (the T of the current phase)."""
        return self.roles()[1]

    def T_old(self):
        return self.roles()[0]

    def step_args(self, phase=None):
        To, T, Tn = self.roles(phase)
        arrs = (_ptr_array(T), _ptr_array(To), _ptr_array(Tn))
        self._keep.append(arrs)
        P = lambda a: ctypes.cast(a, ctypes.POINTER(c_void_p))
        return _lib.FctStepArgs(
            ntracers=self.ntracers, dt=self.dt, uv=self.uv.data_ptr(), w_e=self.w_e.data_ptr(),
            helem=self.helem.data_ptr(), hnode=self.hnode.data_ptr(),
            hnode_new=self.hnode_new.data_ptr(), T=P(arrs[0]), T_old=P(arrs[1]),
            T_new=P(arrs[2]), w_i=None, Z3d=None, zbar3=None)

    def step_alone(self, stream=None, path=_lib.FCT_PATH_FUSED):
        """One step of this instance through the plain entry point (fct_forward_ex)."""
        torch = _torch()
        args = self.step_args()
        s = stream if stream is not None else torch.cuda.current_stream(self.plan.device)
        _lib.check(self.plan.lib.fct_forward_ex(self.plan.ptr, ctypes.byref(args), int(path),
                                                c_void_p(s.cuda_stream)), "fct_forward_ex")
        self.phase = (self.phase + 1) % 3
        self._keep.clear()


class FctBatch:
    """One fused FCT step of every instance per launch set; see module docstring."""

    def __init__(self, instances):
        self.lib = _lib.load()
        self.instances = list(instances)
        n = len(self.instances)
        plans = (c_void_p * n)(*[inst.plan.ptr.value for inst in self.instances])
        ptr = c_void_p()
        _lib.check(self.lib.fct_batch_create(plans, n, ctypes.byref(ptr)), "fct_batch_create")
        self.ptr = ptr
        self._bound = [None] * _lib.FCT_BATCH_SLOTS   # phase offset bound in each slot
        self.tuning = {}

    def close(self):
        if self.ptr:
            self.lib.fct_batch_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, knob: int, value: int):
        _lib.check(self.lib.fct_batch_set_tuning(self.ptr, int(knob), int(value)),
                   "fct_batch_set_tuning")
        self.tuning[int(knob)] = int(value)

    def _bind(self, slot: int, offset: int):
        args = (_lib.FctStepArgs * len(self.instances))(
            *[inst.step_args((inst.phase + offset) % 3) for inst in self.instances])
        _lib.check(self.lib.fct_batch_bind(self.ptr, slot, args), "fct_batch_bind")
        for inst in self.instances:
            inst._keep.clear()
        self._bound[slot] = offset

    def prepare(self):
        """Bind the three rotation phases (slots 0-2) relative to the instances' current
        phases. Re-run after changing an instance's buffers or phase from outside."""
        for r in range(3):
            self._bind(r, r)
        self._base = [inst.phase for inst in self.instances]
        self._r = 0

    def step(self, *, graph: bool = True, stream=None):
        """Advance every instance one step (stream-ordered; no synchronisation)."""
        torch = _torch()
        if self._bound[0] is None:
            self.prepare()
        s = stream if stream is not None else torch.cuda.current_stream(
            self.instances[0].plan.device)
        fn = self.lib.fct_batch_graph_launch if graph else self.lib.fct_batch_launch
        _lib.check(fn(self.ptr, self._r, c_void_p(s.cuda_stream)), "fct_batch launch")
        self._r = (self._r + 1) % 3
        for inst in self.instances:
            inst.phase = (inst.phase + 1) % 3

    def updates_per_step(self) -> int:
        return sum(inst.updates for inst in self.instances)
