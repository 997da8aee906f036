"""Serving FCT requests with continuous batching.

A request asks for ``nsteps`` FCT steps of ``K`` tracers on an icosahedral mesh
(level, nl) with seeded fields, optionally followed by the adjoint of the loss
``sum over wet surface nodes of (T_final - T_target)^2`` (gradients with respect to the
initial tracers and the horizontal velocity). There is no counterpart in the
reference, which advances one state per call (``fesom_jax/tracer_adv.py:532-578``).

``Server`` keeps every admitted request's state resident on the GPU and, on every
iteration, advances all active requests by one step with one batched launch set
(``FctBatch``). A request is admitted (in arrival order) when a slot is free and its
estimated device memory fits the HBM budget; requests arriving while others run join
the next iteration. A finished request's result is downloaded and its memory and
plan returned to a pool. ``fifo_loop`` is the plain per-request loop it is compared
with: one request at a time, each with its own plan, run and download, synchronised.
"""

from __future__ import annotations

import collections
import itertools
import time
from dataclasses import dataclass, field

import numpy as np

from .perfmodel import updates_per_step


@dataclass
class RequestSpec:
    rid: int
    arrival: float          # seconds after the stream starts
    level: int
    nl: int
    K: int
    nsteps: int
    adjoint: bool
    seed: int
    # test hook: "nan" puts a NaN into the request's first tracer, "fail" makes its
    # setup fail (an argument the library refuses); None for ordinary requests
    inject: str | None = None


def make_stream(rate, duration, *, levels=(4, 7), nls=(47, 70), ks=(2, 3, 4), steps=(1, 24),
                p_adjoint=0.3, seed=0):
    """Poisson arrivals at ``rate`` per second for ``duration`` seconds. Levels are
    equally likely (log-uniform over the node counts, which grow 4x per level)."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        out.append(RequestSpec(rid=len(out), arrival=t,
                               level=int(rng.integers(levels[0], levels[1] + 1)),
                               nl=int(rng.choice(nls)), K=int(rng.choice(ks)),
                               nsteps=int(rng.integers(steps[0], steps[1] + 1)),
                               adjoint=bool(rng.random() < p_adjoint),
                               seed=int(rng.integers(1 << 30))))
    return out


class MeshCache:
    """Generated meshes by (level, nl); geometry is shared, plans are not."""

    def __init__(self, seed=7):
        self.seed = seed
        self.meshes = {}

    def get(self, level, nl):
        from .meshgen import icosahedral_mesh
        key = (level, nl)
        if key not in self.meshes:
            self.meshes[key] = icosahedral_mesh(level, nl, seed=self.seed + 100 * level + nl)
        return self.meshes[key]


def estimate_table_bytes(mesh):
    """Device bytes of a mesh's tables (topology, geometry, stencil tables; shared by
    every plan of the mesh the server clones from one another), an upper estimate."""
    N, E, D, L = mesh.nod2D, mesh.elem2D, mesh.edge2D, mesh.nl
    return int(300 * N + 200 * E + 160 * D + 8 * 4 * N * L)   # + area, areasvol, pairs


def estimate_work_bytes(mesh, K, adjoint):
    """Device bytes of a plan's own workspace: the fused path's for ``K`` tracers and,
    once it has run an adjoint, the phased and adjoint workspace (an upper estimate;
    the server replaces it by the plan's own count once the plan exists)."""
    N, E, D, L = mesh.nod2D, mesh.elem2D, mesh.edge2D, mesh.nl
    NL, EL, DL = N * L, E * L, D * L
    total = 8 * (DL + K * (6 * NL + DL))
    if adjoint:
        # phased tape, cotangent scratch, element -> edge lists
        total += 8 * (12 * NL + 2 * EL + DL) + 8 * (20 * NL + 10 * DL + 2 * EL)
        total += 16 * D + 8 * E + 64
    return int(total)


def estimate_plan_bytes(mesh, K, adjoint):
    """Device bytes of a stand-alone plan for ``mesh`` with ``K`` tracers: its tables
    and its workspace."""
    return estimate_table_bytes(mesh) + estimate_work_bytes(mesh, K, adjoint)


def estimate_field_bytes(mesh, K, nsteps, adjoint):
    """Device bytes of a request's own fields: flow, three buffers per tracer and
    (adjoint) the target, the checkpointed states, one recomputed segment of them and
    the cotangents a step's backward holds (g_T, g_T_old in and out, g_uv twice)."""
    N, E, L = mesh.nod2D, mesh.elem2D, mesh.nl
    NL, EL = N * L, E * L
    total = 8 * (3 * EL + 3 * NL + 3 * K * NL)
    if adjoint:
        # checkpoints and one recomputed segment of states (T, T_old)
        total += 8 * (2 * checkpoint_stride(nsteps) + 1) * 2 * K * NL
        total += 8 * (5 * K * NL + 4 * EL)
    return int(total)


def checkpoint_stride(nsteps):
    """Steps between stored states of an adjoint request: about sqrt(nsteps), so the
    stored checkpoints and the one segment recomputed at a time both hold about
    sqrt(nsteps) states."""
    return max(1, int(np.ceil(np.sqrt(nsteps))))


def estimate_base_bytes(mesh):
    """Device bytes of a mesh's base inputs (flow, layer thicknesses and four tracers
    with their previous values), built once per mesh and shared by its requests."""
    NL, EL = mesh.nod2D * mesh.nl, mesh.elem2D * mesh.nl
    return int(8 * (3 * EL + 3 * NL + 8 * NL))


def estimate_bytes(mesh, K, nsteps, adjoint):
    """Device bytes a request needs when it brings its own plan."""
    return estimate_plan_bytes(mesh, K, adjoint) + estimate_field_bytes(mesh, K, nsteps, adjoint)


@dataclass
class _Active:
    spec: RequestSpec
    inst: object
    mesh: object
    case_extra: dict
    remaining: int
    bytes: int
    t_admit: float
    setup_s: float
    # adjoint requests: states (T, T_old) at every `stride`-th step (checkpoints); the
    # backward recomputes the states in between, one segment at a time
    ckpt: dict = field(default_factory=dict)
    stride: int = 1
    done_steps: int = 0


def staged_to_host(src, stage, wait):
    """A copy of the 1-D tensor ``src`` in host memory, made through ``stage`` chunk by
    chunk: each chunk goes into the staging tensor (pinned host memory on a GPU, where
    the copy engine fills it at full speed), ``wait()`` lets that copy finish, and the
    host copies the chunk out. Returns a CPU tensor of ``src``'s dtype."""
    import torch
    out = torch.empty(src.numel(), dtype=src.dtype)
    cs = stage.numel()
    for off in range(0, src.numel(), cs):
        m = min(cs, src.numel() - off)
        stage[:m].copy_(src[off:off + m], non_blocking=True)
        wait()
        out[off:off + m].copy_(stage[:m])
    return out


class Server:
    """Continuous-batching loop; see module docstring.

    Memory: the budget covers every request's own fields, the base inputs of the
    meshes in use and every plan the server holds, active or idle in the pool, at the
    plan's own byte count. The adjoint workspace lives in one adjoint plan per mesh
    (the adjoints run one after another on the finishing stream), counted from the
    first admission that asks for it. Idle plans and unused base inputs are released
    when an admission would not fit otherwise.

    Admission is in arrival order, except that while the oldest waiting request does
    not fit (for at most ``backfill_s`` seconds) a later arrival that fits as things
    stand may start first.

    Failures stay with their request: a request whose setup raises, whose batched step
    fails (found by stepping the batch's instances one at a time), or whose result or
    gradients are not finite is reported with an ``"error"`` entry and
    leaves; the others carry on."""

    def __init__(self, device, *, budget_bytes=200 << 30, max_active=64, meshes=None,
                 dt=2700.0, chunk=4, backfill_s=1.0, lookahead=32):
        self.device = device
        self.chunk = int(chunk)        # steps per iteration while requests are waiting
        self.backfill_s = float(backfill_s)   # longest wait of the oldest request that
        self.lookahead = int(lookahead)       # later arrivals (this many) may overtake
        self.pinned_bytes = 1 << 30           # results up to this size return pinned
        self.stage_elems = 1 << 25            # larger ones through a 256 MiB pinned stage
        self._stage = None
        self._dl_stream = None
        self.budget = int(budget_bytes)
        self.max_active = int(max_active)
        self.meshes = meshes or MeshCache()
        self.dt = dt
        self.pool = collections.defaultdict(list)   # (level, nl, K) -> idle plans
        self.plan_bytes = {}                        # id(plan) -> its own workspace bytes
        self.table_sets = {}                        # table id -> [bytes, ids of live plans]
        self.plan_table = {}                        # id(plan) -> its table id
        self.adj_plans = {}                         # (level, nl) -> the mesh's adjoint plan
        self.adj_ready = set()                      # ids of adjoint plans with workspace
        self.base_bytes = {}                        # (level, nl) -> base input bytes
        self.results = {}
        self.latency = {}
        self.setup = {}
        self.errors = {}
        self.active = []
        self.peak_bytes = 0
        self.peak_detail = None
        self.over_budget = None     # snapshot of the first moment held > budget, if any
        self._admissions = collections.deque(maxlen=8)
        self.device_peak = 0        # device memory in use (all users), highest seen
        self.timing = collections.Counter()         # seconds per activity (host clock)
        self.counts = collections.Counter()         # plans built/closed, base inputs built/dropped
        self._finishing = []

    # -- memory accounting ---------------------------------------------------------
    def held_bytes(self, active):
        """Bytes the server holds: the fields of the active and the finishing requests,
        the meshes' base inputs, every live plan's workspace and each mesh's tables
        once (its plans share them)."""
        return (sum(a.bytes for a in active) + sum(f[0].bytes for f in self._finishing)
                + sum(self.plan_bytes.values()) + sum(b for b, _ in self.table_sets.values())
                + sum(self.base_bytes.values()))

    def _live_plan(self, level, nl):
        """A live plan of mesh (level, nl) held by the server, or None."""
        for (lv, n, _k), ps in self.pool.items():
            if lv == level and n == nl and ps:
                return ps[-1]
        if (level, nl) in self.adj_plans:
            return self.adj_plans[(level, nl)]
        for a in list(self.active) + [f[0] for f in self._finishing]:
            if a.spec.level == level and a.spec.nl == nl:
                return a.inst.plan
        return None

    def _need(self, s):
        """Bytes an admission of ``s`` adds: its fields, its mesh's base inputs the first
        time, a plan unless one is idle, and the adjoint workspace if its plan has not
        run an adjoint yet."""
        m = self.meshes.get(s.level, s.nl)
        key = (s.level, s.nl, s.K)
        need = estimate_field_bytes(m, s.K, s.nsteps, s.adjoint)
        if (s.level, s.nl) not in self.base_bytes:
            need += estimate_base_bytes(m)
        if s.adjoint and (s.level, s.nl) not in self.adj_plans:
            need += estimate_work_bytes(m, 1, True)
        if not self.pool[key]:
            need += estimate_work_bytes(m, s.K, False)
        if self._live_plan(s.level, s.nl) is None:
            need += estimate_table_bytes(m)
        return need

    def _note_peak(self, active):
        import torch
        free, total = torch.cuda.mem_get_info(self.device)
        self.device_peak = max(self.device_peak, total - free)
        held = self.held_bytes(active)
        if held > self.budget and self.over_budget is None:
            self.over_budget = self._snapshot(active, held)
        if held > self.peak_bytes:
            self.peak_bytes = held
            self.peak_detail = {
                "fields": sum(a.bytes for a in active), "active": len(active),
                "plans": {k: sum(self.plan_bytes.get(id(p), 0) for p in ps)
                          for k, ps in self.pool.items() if ps},
                "active_plans": sum(self.plan_bytes.get(id(a.inst.plan), 0) for a in active),
                "all_plans": sum(self.plan_bytes.values()), "base": dict(self.base_bytes)}

    def _snapshot(self, active, held):
        """What the server holds, part by part, and its last admissions (need, held
        before and after): the record of the first moment the held bytes passed the
        budget."""
        label = {}
        for key, ps in self.pool.items():
            for p in ps:
                label[id(p)] = f"idle {key}"
        for key, p in self.adj_plans.items():
            label[id(p)] = f"adjoint {key}"
        for a in active:
            label[id(a.inst.plan)] = f"active {a.spec.rid}"
        for f in self._finishing:
            label[id(f[0].inst.plan)] = f"finishing {f[0].spec.rid}"
        return {"held": int(held),
                "active": {a.spec.rid: a.bytes for a in active},
                "finishing": {f[0].spec.rid: f[0].bytes for f in self._finishing},
                "plans": {label.get(k, f"unlabelled {k}"): v for k, v in self.plan_bytes.items()},
                "tables": {str(t): b for t, (b, _) in self.table_sets.items()},
                "base": {str(k): v for k, v in self.base_bytes.items()},
                "admissions": list(self._admissions)}

    def _refresh(self, plan):
        self.plan_bytes[id(plan)] = plan.own_bytes()
        tid, tbytes = plan.tables()
        self.plan_table[id(plan)] = tid
        self.table_sets.setdefault(tid, [tbytes, set()])[1].add(id(plan))

    def _forget(self, plan):
        """Drop a closed plan from the accounting (its tables with their last plan)."""
        del self.plan_bytes[id(plan)]
        tid = self.plan_table.pop(id(plan))
        users = self.table_sets[tid][1]
        users.discard(id(plan))
        if not users:
            del self.table_sets[tid]

    def _release_idle(self, need, active, busy=(), keep=None, keep_base=None, keep_adj=None):
        """Close idle plans (largest first, never ``keep``: the plan the admission that
        asks will reuse, whose bytes ``need`` leaves out), then drop the base inputs of
        meshes no active or finishing request uses, nor the asking admission
        (``keep_base``; they are rebuilt when needed), until ``need`` more bytes fit;
        returns whether they do. ``need`` may be a function of the server's state,
        evaluated after every release: closing the last live plan of the asking
        admission's mesh makes that admission build the mesh's tables again."""
        need_now = need if callable(need) else (lambda: need)
        # an adjoint plan is in use while a request of its mesh that asks for the adjoint
        # is active or finishing; it is kept for an asking admission that runs an adjoint
        # (``keep_adj``, its mesh), while a forward-only one may release it
        adj_use = {(a.spec.level, a.spec.nl) for a in list(active) + list(busy) if a.spec.adjoint}
        if keep_adj is not None:
            adj_use.add(keep_adj)
        idle = sorted([(self.plan_bytes[id(p)], k, p) for k, ps in self.pool.items() for p in ps
                       if p is not keep] +
                      [(self.plan_bytes[id(p)], k, p) for k, p in self.adj_plans.items()
                       if k not in adj_use], key=lambda x: -x[0])
        for nbytes, key, plan in idle:
            if self.held_bytes(active) + need_now() <= self.budget:
                return True
            if key in self.adj_plans and self.adj_plans[key] is plan:
                del self.adj_plans[key]
                self.adj_ready.discard(id(plan))
            else:
                self.pool[key].remove(plan)
            self._forget(plan)
            tt = time.perf_counter()
            plan.close()          # frees device memory: waits for the device's queued work
            self.timing["release"] += time.perf_counter() - tt
            self.counts["plans_closed"] += 1
        in_use = {(a.spec.level, a.spec.nl) for a in list(active) + list(busy)}
        if keep_base is not None:
            in_use.add(keep_base)
        for key in sorted(self.base_bytes, key=lambda k: -self.base_bytes[k]):
            if self.held_bytes(active) + need_now() <= self.budget:
                break
            if key in in_use:
                continue
            tt = time.perf_counter()
            _drop_base_case(self.meshes.get(*key), self.device, self.dt)
            self.timing["release"] += time.perf_counter() - tt
            del self.base_bytes[key]
            self.counts["base_dropped"] += 1
        return self.held_bytes(active) + need_now() <= self.budget

    # -- request life cycle --------------------------------------------------------
    def _plan(self, mesh, spec):
        from .tracer_adv import FctPlan
        import torch
        key = (spec.level, spec.nl, spec.K)
        if self.pool[key]:
            return self.pool[key].pop()
        # memory torch caches but no tensor uses would otherwise stay out of the plan's
        # reach
        torch.cuda.empty_cache()
        t = time.perf_counter()
        share = self._live_plan(spec.level, spec.nl)
        plan = FctPlan(mesh, self.device, max_tracers=spec.K, share=share)
        self._refresh(plan)
        self.timing["plan_create" if share is None else "plan_clone"] += time.perf_counter() - t
        self.counts["plans_built" if share is None else "plans_cloned"] += 1
        return plan

    def _adjoint_plan(self, mesh, spec, share=None):
        """The mesh's adjoint plan (one tracer of fused workspace: it only runs the
        backward, one request after another on the finishing stream); its adjoint
        workspace is counted from admission on. It shares the tables of ``share`` (the
        admitted request's own plan, not yet among the active ones) or of another live
        plan of the mesh."""
        from .tracer_adv import FctPlan
        import torch
        key = (spec.level, spec.nl)
        if key not in self.adj_plans:
            torch.cuda.empty_cache()
            t = time.perf_counter()
            plan = FctPlan(mesh, self.device, max_tracers=1,
                           share=share if share is not None else self._live_plan(spec.level, spec.nl))
            self.adj_plans[key] = plan
            self._refresh(plan)
            # the adjoint workspace is allocated by the first backward: count it now
            self.plan_bytes[id(plan)] = estimate_work_bytes(mesh, 1, True)
            self.timing["plan_clone"] += time.perf_counter() - t
            self.counts["adjoint_plans_built"] += 1
        return self.adj_plans[key]

    def _start(self, spec, now):
        from .batch import Instance
        t = time.perf_counter()
        mesh = self.meshes.get(spec.level, spec.nl)
        self.base_bytes.setdefault((spec.level, spec.nl), estimate_base_bytes(mesh))
        tc = time.perf_counter()
        if not _has_base_case(mesh, self.device, self.dt):
            _base_case(mesh, self.device, self.dt)
            self.timing["base_inputs"] += time.perf_counter() - tc
            self.counts["base_built"] += 1
            tc = time.perf_counter()
        case = request_case(mesh, spec, self.device, self.dt)
        self.timing["inputs"] += time.perf_counter() - tc
        plan = self._plan(mesh, spec)
        try:
            inst = Instance.from_case(mesh, case, self.device, plan=plan)
        except Exception:
            self.pool[(spec.level, spec.nl, spec.K)].append(plan)
            raise
        extra = {"uv": case["uv"], "target": case.get("target")}
        import torch
        # the request's inputs are ready once the main stream's queued work is; the
        # finishing stream's adjoints keep running meanwhile
        torch.cuda.current_stream(self.device).synchronize()
        nbytes = estimate_field_bytes(mesh, spec.K, spec.nsteps, spec.adjoint)
        if spec.adjoint:
            self._adjoint_plan(mesh, spec, share=plan)
        a = _Active(spec=spec, inst=inst, mesh=mesh, case_extra=extra, remaining=spec.nsteps,
                    bytes=nbytes, t_admit=now, setup_s=time.perf_counter() - t)
        if spec.adjoint:
            a.stride = checkpoint_stride(spec.nsteps)
            a.ckpt[0] = ([x.clone() for x in inst.T()], [x.clone() for x in inst.T_old()])
        return a

    def _finish_async(self, a, side):
        """Queue the request's finishing work (adjoint on the request's own plan,
        finite check, download into pinned host memory) on the side stream after the
        steps queued so far; returns the event that marks its completion, the host
        buffers and the device tensors they come from."""
        import torch
        mkey = (a.spec.level, a.spec.nl)
        adj_plan = self.adj_plans.get(mkey)
        if a.spec.adjoint and adj_plan is not None and id(adj_plan) not in self.adj_ready:
            # the plan's first adjoint allocates its workspace: hand back what torch
            # caches but no tensor uses
            torch.cuda.empty_cache()
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(self.device))
        side.wait_event(ev)
        with torch.cuda.stream(side):
            out = {"T": [x for x in a.inst.T()]}
            if a.spec.adjoint:
                out.update(adjoint_of_rollout(a, self.dt, to_host=False, adj_plan=adj_plan))
                self.adj_ready.add(id(adj_plan))
            # one reduction per field (a NaN or an infinity anywhere makes the sum
            # non-finite), so the check allocates nothing the size of a field
            bad = torch.zeros((), dtype=torch.bool, device=self.device)
            for x in out["T"] + out.get("g_T0", []):
                bad |= ~torch.isfinite(x.sum())
            if "g_uv" in out:
                bad |= ~torch.isfinite(out["g_uv"].sum())
            host = {}
            nbytes = sum(x.numel() * x.element_size()
                         for v in out.values() for x in (v if isinstance(v, list) else [v]))
            if nbytes <= self.pinned_bytes:
                # small results come back through pinned buffers on the finishing stream
                for k, v in out.items():
                    vs = v if isinstance(v, list) else [v]
                    hs = [torch.empty(x.shape, dtype=x.dtype, pin_memory=True) for x in vs]
                    for h, x in zip(hs, vs):
                        h.copy_(x, non_blocking=True)
                    host[k] = hs if isinstance(v, list) else hs[0]
            else:
                # large ones are copied out once finished (pinned host memory is scarce)
                host["_deferred"] = True
            host["_bad"] = torch.empty((), dtype=torch.bool, pin_memory=True)
            host["_bad"].copy_(bad, non_blocking=True)
            done = torch.cuda.Event()
            done.record(side)
        # the checkpoints and the request's buffers stay referenced (through `a`) until
        # `done` has completed, so the allocator cannot hand them out earlier
        return done, host, out

    def _fail(self, spec, t0, msg, log=None):
        self.errors[spec.rid] = msg
        self.results[spec.rid] = {"error": msg}
        self.latency[spec.rid] = time.perf_counter() - t0 - spec.arrival
        if log:
            log(f"req {spec.rid} failed: {msg}")

    def _step_isolating(self, batch, active, graph, t0, log):
        """One batched step; if it fails, step the instances one at a time so that only
        the failing ones leave (their state is not advanced further)."""
        try:
            batch.step(graph=graph)
            return list(active), []
        except RuntimeError as exc:
            first = str(exc)
        ok, bad = [], []
        for a in active:
            try:
                a.inst.step_alone()
                ok.append(a)
            except RuntimeError as exc:
                bad.append((a, f"step failed: {exc} (batched launch: {first})"))
        return ok, bad

    # -- the loop --------------------------------------------------------------------
    def run(self, specs, *, graph=False, log=None):
        """Serve ``specs`` in real time (arrival offsets from the call). Returns a
        summary dict; per-request results, latencies and setup times are kept on the
        server."""
        import torch
        from .batch import FctBatch
        pending = collections.deque(sorted(specs, key=lambda s: s.arrival))
        active = []
        self.active = active
        batch, changed, retired = None, False, []
        finishing = self._finishing
        finishing.clear()
        side = torch.cuda.Stream(self.device)      # default priority
        t0 = time.perf_counter()
        busy = 0.0
        updates = 0
        while pending or active or finishing:
            now = time.perf_counter() - t0
            while (pending and pending[0].arrival <= now and len(active) < self.max_active):
                s = pending[0]
                ps = self.pool[(s.level, s.nl, s.K)]
                fits = self._release_idle(lambda s=s: self._need(s), active,
                                          [f[0] for f in finishing],
                                          keep=ps[-1] if ps else None,
                                          keep_base=(s.level, s.nl),
                                          keep_adj=(s.level, s.nl) if s.adjoint else None)
                if not fits and not (active or finishing):
                    # larger than the budget with everything else released: runs alone
                    self.counts["admitted_alone_over_budget"] += 1
                if not fits and (active or finishing):
                    # the oldest request waits for memory; until it has waited
                    # `backfill_s`, a later arrival that fits as things stand goes ahead
                    s = None
                    if now - pending[0].arrival < self.backfill_s:
                        held = self.held_bytes(active)
                        for c in itertools.islice(pending, 1, 1 + self.lookahead):
                            if c.arrival > now:
                                break
                            if held + self._need(c) <= self.budget:
                                s = c
                                break
                    if s is None:
                        break
                    self.counts["backfilled"] += 1
                pending.remove(s)
                need_s, held_s = self._need(s), self.held_bytes(active)
                try:
                    a = self._start(s, now)
                except Exception as exc:          # setup failure stays with its request
                    self._fail(s, t0, f"setup failed: {exc}", log)
                    continue
                self.setup[s.rid] = a.setup_s
                active.append(a)
                self._admissions.append((s.rid, s.level, s.nl, s.K, bool(s.adjoint), need_s,
                                         held_s, self.held_bytes(active)))
                self._note_peak(active)
                changed = True
            if not active:
                if finishing:
                    # nothing to step: wait for the oldest finishing request
                    tt = time.perf_counter()
                    finishing[0][1].synchronize()
                    self.timing["wait_finishing"] += time.perf_counter() - tt
                elif pending:
                    wait = pending[0].arrival - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
            else:
                tb = time.perf_counter()
                if changed:
                    tt = time.perf_counter()
                    if batch is not None:
                        # closed once the launches queued so far have run
                        ev = torch.cuda.Event()
                        ev.record(torch.cuda.current_stream(self.device))
                        retired.append((batch, ev))
                    try:
                        batch = FctBatch([a.inst for a in active])
                        batch.prepare()
                    except (RuntimeError, MemoryError):
                        # out of device memory for the batch's tables: free what the
                        # retired batches and torch's cache hold, then once more
                        torch.cuda.synchronize(self.device)
                        for old, _ in retired:
                            old.close()
                        retired.clear()
                        torch.cuda.empty_cache()
                        self.counts["batch_retry"] += 1
                        batch = FctBatch([a.inst for a in active])
                        batch.prepare()
                    changed = False
                    self.timing["batch_build"] += time.perf_counter() - tt
                # steps until the next completion; at most `chunk` while requests wait
                nsteps = min(a.remaining for a in active)
                if pending:
                    nsteps = min(nsteps, self.chunk)
                tt = time.perf_counter()
                for _ in range(nsteps):
                    stepped, failed = self._step_isolating(batch, active, graph, t0, log)
                    for a, msg in failed:
                        active.remove(a)
                        self._fail(a.spec, t0, msg, log)
                        self.pool[(a.spec.level, a.spec.nl, a.spec.K)].append(a.inst.plan)
                        changed = True
                    if failed or len(stepped) != len(active):
                        # the batch's instances moved one at a time: rebind it
                        changed = True
                    for a in stepped:
                        a.remaining -= 1
                        a.done_steps += 1
                        updates += a.inst.updates
                        if a.spec.adjoint and a.done_steps % a.stride == 0 and a.remaining:
                            a.ckpt[a.done_steps] = ([x.clone() for x in a.inst.T()],
                                                    [x.clone() for x in a.inst.T_old()])
                    if changed or not active:
                        break
                self.timing["steps_queued"] += time.perf_counter() - tt
                done = [a for a in active if a.remaining == 0]
                for a in done:
                    # the instance leaves the batch; its finishing work runs on the side
                    # stream while the remaining requests keep stepping
                    active.remove(a)
                    changed = True
                    try:
                        tt = time.perf_counter()
                        try:
                            ev, host, keep = self._finish_async(a, side)
                        except (RuntimeError, MemoryError) as exc:
                            if "out of memory" not in str(exc):
                                raise
                            # let the finishing work already queued run and free its
                            # memory, then once more
                            side.synchronize()
                            torch.cuda.empty_cache()
                            self.counts["finish_retry"] += 1
                            ev, host, keep = self._finish_async(a, side)
                        self.timing["finish_queued"] += time.perf_counter() - tt
                    except (RuntimeError, MemoryError) as exc:
                        side.synchronize()
                        free, total = torch.cuda.mem_get_info(self.device)
                        self._fail(a.spec, t0, f"finishing failed: {exc} (icosa-{a.spec.level} "
                                   f"x {a.spec.nl}, K {a.spec.K}, {a.spec.nsteps} steps, "
                                   f"adjoint {a.spec.adjoint}; held "
                                   f"{self.held_bytes(active) / 2**30:.1f} GiB, torch allocated "
                                   f"{torch.cuda.memory_allocated(self.device) / 2**30:.1f} GiB, "
                                   f"device free {free / 2**30:.1f} GiB)", log)
                        self.pool[(a.spec.level, a.spec.nl, a.spec.K)].append(a.inst.plan)
                        self._refresh(a.inst.plan)
                        continue
                    finishing.append((a, ev, host, keep))
                busy += time.perf_counter() - tb
            still = []
            for a, ev, host, keep in finishing:
                if ev.query():
                    bad = bool(host.pop("_bad"))
                    if host.pop("_deferred", False):
                        tt = time.perf_counter()
                        host = {k: ([self._to_host(x) for x in v] if isinstance(v, list)
                                    else self._to_host(v)) for k, v in keep.items()}
                        self.timing["download"] += time.perf_counter() - tt
                    if bad:
                        self._fail(a.spec, t0, "non-finite result", log)
                    else:
                        self.results[a.spec.rid] = {
                            k: ([h.numpy() for h in v] if isinstance(v, list) else v.numpy())
                            for k, v in host.items()}
                        self.latency[a.spec.rid] = time.perf_counter() - t0 - a.spec.arrival
                    mk = (a.spec.level, a.spec.nl)
                    if a.spec.adjoint and mk in self.adj_plans:
                        self._refresh(self.adj_plans[mk])   # its workspace now exists
                    self.pool[(a.spec.level, a.spec.nl, a.spec.K)].append(a.inst.plan)
                    if log and not bad:
                        free, total = torch.cuda.mem_get_info(self.device)
                        log(f"req {a.spec.rid} (icosa-{a.spec.level} x {a.spec.nl}, K "
                            f"{a.spec.K}, {a.spec.nsteps} steps{', adjoint' if a.spec.adjoint else ''})"
                            f" done in {self.latency[a.spec.rid]:.3f}s; held "
                            f"{self.held_bytes(active) / 2**30:.1f} GiB, device in use "
                            f"{(total - free) / 2**30:.1f} GiB")
                else:
                    still.append((a, ev, host, keep))
            finishing[:] = still         # the same list object as self._finishing
            self._note_peak(active)
            done_r = [(b, e) for b, e in retired if e.query()]
            for old, _ in done_r:
                old.close()
            retired[:] = [r for r in retired if r not in done_r]
        torch.cuda.synchronize(self.device)
        for old in [b for b, _ in retired] + ([batch] if batch is not None else []):
            old.close()
        wall = time.perf_counter() - t0
        out = summarize([s for s in specs if s.rid not in self.errors], self.latency,
                        self.setup, wall, busy, self.meshes)
        out["failed"] = len(self.errors)
        out["peak_bytes"] = int(self.peak_bytes)
        out["budget_bytes"] = self.budget
        if self.over_budget is not None:
            out["over_budget"] = self.over_budget
        if self.peak_detail is not None:
            d = dict(self.peak_detail)
            d["plans"] = {str(k): v for k, v in d["plans"].items()}
            d["base"] = {str(k): v for k, v in d["base"].items()}
            out["peak_detail"] = d
        out["device_peak_bytes"] = int(self.device_peak)
        out["timing_s"] = {k: round(v, 3) for k, v in self.timing.items()}
        out["counts"] = dict(self.counts)
        return out

    def _to_host(self, x):
        """A finished large result in host memory, through the server's pinned staging
        buffer on a stream of its own (the steps queued on the main stream and the
        adjoints on the finishing stream are not waited for)."""
        import torch
        if self._stage is None:
            self._stage = torch.empty(self.stage_elems, dtype=torch.float64, pin_memory=True)
            self._dl_stream = torch.cuda.Stream(self.device)      # default priority
        if x.dtype != torch.float64:
            return x.cpu()
        with torch.cuda.stream(self._dl_stream):
            # a view for the contiguous results; any copy is made on this stream too
            src = x.reshape(-1)
            out = staged_to_host(src, self._stage, self._dl_stream.synchronize)
        return out.reshape(x.shape)

    def close(self):
        """Release every pooled plan."""
        for ps in self.pool.values():
            for p in ps:
                p.close()
        for p in self.adj_plans.values():
            p.close()
        self.pool.clear()
        self.adj_plans.clear()
        self.adj_ready.clear()
        self._stage = None
        self.plan_bytes.clear()
        self.table_sets.clear()
        self.plan_table.clear()


_BASE = {}


def _base_case(mesh, device, dt):
    """Seeded base inputs of a mesh (the configs' recipe), built once per mesh. An
    entry goes with its mesh (so a later mesh that reuses the id never finds it)."""
    import weakref
    from .inputs import make_case_device
    key = (id(mesh), str(device), dt)
    hit = _BASE.get(key)
    if hit is None or hit[0]() is not mesh:
        _BASE[key] = (weakref.ref(mesh), make_case_device(mesh, device, seed=1234, dt=dt,
                                                          tracers=("T", "S", "T", "S")))
        weakref.finalize(mesh, _BASE.pop, key, None)
    return _BASE[key][1]


def _has_base_case(mesh, device, dt):
    hit = _BASE.get((id(mesh), str(device), dt))
    return hit is not None and hit[0]() is mesh


def _drop_base_case(mesh, device, dt):
    _BASE.pop((id(mesh), str(device), dt), None)


def request_case(mesh, spec, device, dt):
    """The request's inputs on the device, standing in for its payload: the mesh's base
    case with the velocity (and w, so continuity holds) scaled by a seeded factor in
    [0.8, 1.2] and seeded noise (0.05 amplitude) on the wet lanes of each tracer;
    T_old = T. Adjoint requests also get a target T + 0.2 N(0, 1)."""
    import torch
    base = _base_case(mesh, device, dt)
    g = torch.Generator(device=device).manual_seed(spec.seed)
    f64 = dict(dtype=torch.float64, device=device)
    nlm = torch.from_numpy(mesh.node_layer_mask).to(device)
    zero = torch.zeros((), **f64)
    fac = 0.8 + 0.4 * float(torch.rand((), generator=g, **f64))
    noise = lambda amp, t: torch.where(nlm, amp * torch.randn(t.shape, generator=g, **f64), zero)
    Ts = [base["T"][j] + noise(0.05, base["T"][j]) for j in range(spec.K)]
    if spec.inject == "nan":
        Ts[0][0, 0] = float("nan")
    if spec.inject == "fail":
        # a tracer of the wrong shape: the library refuses the instance
        Ts[0] = Ts[0][:, :-1]
    case = dict(uv=base["uv"] * fac, w_e=base["w_e"] * fac, helem=base["helem"],
                hnode=base["hnode"], hnode_new=base["hnode_new"], T=Ts,
                T_old=[t.clone() for t in Ts], dt=dt)
    if spec.adjoint:
        case["target"] = [t + noise(0.2, t) for t in Ts]
    return case


def adjoint_of_rollout(a, dt, to_host=True, adj_plan=None):
    """Gradients of sum_surface (T_final - T_target)^2 with respect to the initial
    tracers (T and T_old of step 0) and uv. The backward runs segment by segment from
    the last checkpoint to the first: each segment's states are recomputed from its
    checkpoint with the request's own plan (the same bits as the steps served), then
    the step adjoints run back through them on ``adj_plan`` (a plan of the same mesh
    holding the adjoint workspace; the request's own plan by default)."""
    import torch
    from .adjoint import advect_fct_vjp
    from .tracer_adv import advect_fct
    inst, mesh = a.inst, a.mesh
    n = a.spec.nsteps
    surf = torch.from_numpy(mesh.node_layer_mask[:, 0]).to(inst.uv.device)
    fin = inst.T()
    g = []
    for Tf, tg in zip(fin, a.case_extra["target"]):
        gj = torch.zeros_like(Tf)
        gj[:, 0] = torch.where(surf, 2.0 * (Tf[:, 0] - tg[:, 0]), torch.zeros_like(Tf[:, 0]))
        g.append(gj)
    gTo = [torch.zeros_like(x) for x in g]
    g_uv = torch.zeros_like(inst.uv)
    flow = (inst.uv, inst.w_e, inst.helem, inst.hnode, inst.hnode_new)
    for s0 in range((n - 1) // a.stride * a.stride, -1, -a.stride):
        s1 = min(s0 + a.stride, n)
        T, To = a.ckpt[s0]
        states = [(T, To)]
        for _ in range(s0, s1 - 1):
            Tn = [o[0] for o in advect_fct(mesh, *flow, T, To, dt=dt, plan=inst.plan)]
            To, T = T, Tn
            states.append((T, To))
        for T, To in reversed(states):
            gT, gTo_s, guv, _ = advect_fct_vjp(mesh, *flow, T, To, g, dt=dt, need_w=False,
                                               plan=adj_plan or inst.plan)
            g_uv += guv
            g = [x + y for x, y in zip(gT, gTo)]
            gTo = gTo_s
    if not to_host:
        return {"g_T0": g, "g_T_old0": gTo, "g_uv": g_uv}
    return {"g_T0": [x.cpu().numpy() for x in g], "g_T_old0": [x.cpu().numpy() for x in gTo],
            "g_uv": g_uv.cpu().numpy()}


def fifo_loop(specs, device, *, meshes=None, dt=2700.0, budget_bytes=200 << 30):
    """The plain per-request loop: requests in arrival order, one at a time; each
    builds its plan and inputs, runs its steps with the plain entry point, downloads
    its result and synchronises before the next one starts. Plans are kept per (mesh,
    tracers) for later requests and base inputs per mesh while they, with the request's
    own fields, stay within ``budget_bytes`` (the least recently used plans are closed
    beyond it, then the base inputs of other meshes dropped)."""
    import torch
    from .batch import Instance
    from .tracer_adv import FctPlan
    meshes = meshes or MeshCache()
    lat, setup, results = {}, {}, {}
    plans = {}
    adjoint_ready = set()          # ids of plans that hold the adjoint workspace
    bases = {}                     # (level, nl) -> base input bytes, least recent first
    counts, timing = collections.Counter(), collections.Counter()
    t0 = time.perf_counter()
    busy = 0.0
    for spec in sorted(specs, key=lambda s: s.arrival):
        now = time.perf_counter() - t0
        if spec.arrival > now:
            time.sleep(spec.arrival - now)
        tb = time.perf_counter()
        mesh = meshes.get(spec.level, spec.nl)
        key, mkey = (spec.level, spec.nl, spec.K), (spec.level, spec.nl)
        need = estimate_field_bytes(mesh, spec.K, spec.nsteps, spec.adjoint)
        if mkey not in bases:
            need += estimate_base_bytes(mesh)
        if key not in plans:
            need += estimate_plan_bytes(mesh, spec.K, spec.adjoint)
        elif spec.adjoint and id(plans[key]) not in adjoint_ready:
            need += (estimate_plan_bytes(mesh, spec.K, True)
                     - estimate_plan_bytes(mesh, spec.K, False))
        held = lambda: sum(p.device_bytes() for p in plans.values()) + sum(bases.values())
        for k in [k for k in plans if k != key]:
            if held() + need <= budget_bytes:
                break
            p = plans.pop(k)
            adjoint_ready.discard(id(p))
            p.close()
            counts["plans_closed"] += 1
        for k in [k for k in bases if k != mkey]:
            if held() + need <= budget_bytes:
                break
            _drop_base_case(meshes.get(*k), device, dt)
            del bases[k]
            counts["base_dropped"] += 1
        torch.cuda.empty_cache()
        tc = time.perf_counter()
        if not _has_base_case(mesh, device, dt):
            counts["base_built"] += 1
        bases[mkey] = bases.pop(mkey, estimate_base_bytes(mesh))     # most recent last
        case = request_case(mesh, spec, device, dt)
        timing["inputs"] += time.perf_counter() - tc
        if key not in plans:
            tc = time.perf_counter()
            plans[key] = FctPlan(mesh, device, max_tracers=spec.K)
            timing["plan_create"] += time.perf_counter() - tc
            counts["plans_built"] += 1
        plans[key] = plans.pop(key)          # most recently used last
        inst = Instance.from_case(mesh, case, device, plan=plans[key])
        torch.cuda.synchronize(device)
        setup[spec.rid] = time.perf_counter() - tb
        a = _Active(spec=spec, inst=inst, mesh=mesh,
                    case_extra={"uv": case["uv"], "target": case.get("target")},
                    remaining=spec.nsteps, bytes=0, t_admit=0.0, setup_s=setup[spec.rid])
        if spec.adjoint:
            a.stride = checkpoint_stride(spec.nsteps)
            a.ckpt[0] = ([x.clone() for x in inst.T()], [x.clone() for x in inst.T_old()])
        for i in range(spec.nsteps):
            inst.step_alone()
            if spec.adjoint and (i + 1) % a.stride == 0 and i + 1 < spec.nsteps:
                a.ckpt[i + 1] = ([x.clone() for x in inst.T()], [x.clone() for x in inst.T_old()])
        res = {"T": [x.cpu().numpy() for x in inst.T()]}
        if spec.adjoint:
            res.update(adjoint_of_rollout(a, dt))
            adjoint_ready.add(id(plans[key]))
        torch.cuda.synchronize(device)
        results[spec.rid] = res
        lat[spec.rid] = time.perf_counter() - t0 - spec.arrival
        busy += time.perf_counter() - tb
        del inst, a, case
    wall = time.perf_counter() - t0
    for p in plans.values():
        p.close()
    out = summarize(specs, lat, setup, wall, busy, meshes)
    out["timing_s"] = {k: round(v, 3) for k, v in timing.items()}
    out["counts"] = dict(counts)
    out["results"] = results
    return out


def summarize(specs, lat, setup, wall, busy, meshes):
    """Latency percentiles and forward node-level updates per second of wall time
    (first arrival to last completion), with and without the time spent in
    per-request setup (input materialisation, plans)."""
    done = [s for s in specs if s.rid in lat]
    fwd_updates = sum(updates_per_step(meshes.get(s.level, s.nl), s.K) * s.nsteps for s in done)
    l = np.array([lat[s.rid] for s in done]) if done else np.zeros(1)
    st = float(sum(setup.get(s.rid, 0.0) for s in done))
    return {"requests": len(done), "wall_s": wall, "busy_s": busy,
            "p50_s": float(np.percentile(l, 50)), "p99_s": float(np.percentile(l, 99)),
            "updates": int(fwd_updates), "updates_per_s": fwd_updates / wall,
            "setup_s": st, "updates_per_s_excl_setup": fwd_updates / max(wall - st, 1e-9)}
