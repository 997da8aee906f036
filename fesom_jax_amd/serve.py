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


def estimate_bytes(mesh, K, nsteps, adjoint):
    """Device bytes a request holds: plan topology and fused workspace, fields, and
    (adjoint) the stored trajectory plus the adjoint workspace."""
    N, E, D, L = mesh.nod2D, mesh.elem2D, mesh.edge2D, mesh.nl
    NL, EL, DL = N * L, E * L, D * L
    topo = 200 * N + 200 * E + 150 * D + 8 * 4 * NL     # + area, areasvol and their pairs
    work = 8 * (DL + K * (6 * NL + DL))
    fields = 8 * (3 * EL + 3 * NL + 3 * K * NL)
    total = topo + work + fields
    if adjoint:
        total += 8 * (nsteps + 1) * 2 * K * NL            # trajectory (T, T_old)
        total += 8 * (12 * NL + 2 * EL + DL) + 8 * (20 * NL + 10 * DL + 2 * EL)
    return int(total)


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
    traj: list = field(default_factory=list)


class Server:
    def __init__(self, device, *, budget_bytes=200 << 30, max_active=64, meshes=None,
                 dt=2700.0, chunk=4):
        self.device = device
        self.chunk = int(chunk)        # steps per iteration while requests are waiting
        self.budget = int(budget_bytes)
        self.max_active = int(max_active)
        self.meshes = meshes or MeshCache()
        self.dt = dt
        self.pool = collections.defaultdict(list)   # (level, nl, K) -> idle plans
        self.results = {}
        self.latency = {}
        self.setup = {}
        self.active = []

    # -- request life cycle --------------------------------------------------------
    def _plan(self, mesh, spec):
        from .tracer_adv import FctPlan
        key = (spec.level, spec.nl, spec.K)
        if self.pool[key]:
            return self.pool[key].pop()
        return FctPlan(mesh, self.device, max_tracers=spec.K)

    def _start(self, spec, now):
        from .batch import Instance
        t = time.perf_counter()
        mesh = self.meshes.get(spec.level, spec.nl)
        case = request_case(mesh, spec, self.device, self.dt)
        inst = Instance.from_case(mesh, case, self.device, plan=self._plan(mesh, spec))
        extra = {"uv": case["uv"], "target": case.get("target")}
        import torch
        torch.cuda.synchronize(self.device)
        a = _Active(spec=spec, inst=inst, mesh=mesh, case_extra=extra, remaining=spec.nsteps,
                    bytes=estimate_bytes(mesh, spec.K, spec.nsteps, spec.adjoint),
                    t_admit=now, setup_s=time.perf_counter() - t)
        if spec.adjoint:
            a.traj.append(([x.clone() for x in inst.T()], [x.clone() for x in inst.T_old()]))
        return a

    def _finish_async(self, a, side):
        """Queue the request's finishing work (adjoint, download into pinned host
        memory) on the side stream after the steps queued so far; returns the event
        that marks its completion and the host buffers."""
        import torch
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(self.device))
        side.wait_event(ev)
        with torch.cuda.stream(side):
            out = {"T": [x for x in a.inst.T()]}
            if a.spec.adjoint:
                out.update(adjoint_of_rollout(a, self.dt, to_host=False))
            host = {}
            for k, v in out.items():
                vs = v if isinstance(v, list) else [v]
                hs = [torch.empty(x.shape, dtype=x.dtype, pin_memory=True) for x in vs]
                for h, x in zip(hs, vs):
                    h.copy_(x, non_blocking=True)
                host[k] = hs if isinstance(v, list) else hs[0]
            done = torch.cuda.Event()
            done.record(side)
        # the trajectory and the request's buffers stay referenced (through `a`) until
        # `done` has completed, so the allocator cannot hand them out earlier
        return done, host, out

    # -- the loop --------------------------------------------------------------------
    def run(self, specs, *, graph=False, log=None):
        """Serve ``specs`` in real time (arrival offsets from the call). Returns a
        summary dict; per-request results, latencies and setup times are kept on the
        server."""
        import torch
        from .batch import FctBatch
        pending = collections.deque(sorted(specs, key=lambda s: s.arrival))
        active, used = [], 0
        self.active = active
        batch, changed, retired = None, False, []
        finishing = []
        side = torch.cuda.Stream(self.device)      # default priority
        t0 = time.perf_counter()
        busy = 0.0
        updates = 0
        while pending or active or finishing:
            now = time.perf_counter() - t0
            while (pending and pending[0].arrival <= now and len(active) < self.max_active):
                s = pending[0]
                m = self.meshes.get(s.level, s.nl)
                need = estimate_bytes(m, s.K, s.nsteps, s.adjoint)
                if used + need > self.budget and active:
                    break
                pending.popleft()
                a = self._start(s, now)
                self.setup[s.rid] = a.setup_s
                used += a.bytes
                active.append(a)
                changed = True
            if not active:
                if finishing:
                    # nothing to step: wait for the oldest finishing request
                    finishing[0][1].synchronize()
                elif pending:
                    wait = pending[0].arrival - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
            else:
                tb = time.perf_counter()
                if changed:
                    if batch is not None:
                        retired.append(batch)     # closed once the GPU is idle
                    batch = FctBatch([a.inst for a in active])
                    batch.prepare()
                    changed = False
                # steps until the next completion; at most `chunk` while requests wait
                nsteps = min(a.remaining for a in active)
                if pending:
                    nsteps = min(nsteps, self.chunk)
                for _ in range(nsteps):
                    batch.step(graph=graph)
                    for a in active:
                        a.remaining -= 1
                        updates += a.inst.updates
                        if a.spec.adjoint:
                            a.traj.append(([x.clone() for x in a.inst.T()],
                                           [x.clone() for x in a.inst.T_old()]))
                done = [a for a in active if a.remaining == 0]
                for a in done:
                    # the instance leaves the batch; its finishing work runs on the side
                    # stream while the remaining requests keep stepping
                    ev, host, keep = self._finish_async(a, side)
                    finishing.append((a, ev, host, keep))
                    active.remove(a)
                    changed = True
                busy += time.perf_counter() - tb
            still = []
            for a, ev, host, keep in finishing:
                if ev.query():
                    self.results[a.spec.rid] = {
                        k: ([h.numpy() for h in v] if isinstance(v, list) else v.numpy())
                        for k, v in host.items()}
                    self.latency[a.spec.rid] = time.perf_counter() - t0 - a.spec.arrival
                    used -= a.bytes
                    self.pool[(a.spec.level, a.spec.nl, a.spec.K)].append(a.inst.plan)
                    if log:
                        log(f"req {a.spec.rid} done in {self.latency[a.spec.rid]:.3f}s")
                else:
                    still.append((a, ev, host, keep))
            finishing[:] = still
            if len(retired) > 16 and not finishing and not active:
                for old in retired:
                    old.close()
                retired.clear()
        torch.cuda.synchronize(self.device)
        for old in retired + ([batch] if batch is not None else []):
            old.close()
        wall = time.perf_counter() - t0
        return summarize(specs, self.latency, self.setup, wall, busy, self.meshes)


_BASE = {}


def _base_case(mesh, device, dt):
    """Seeded base inputs of a mesh (the configs' recipe), built once per mesh."""
    from .inputs import make_case_device
    key = (id(mesh), str(device), dt)
    if key not in _BASE:
        _BASE[key] = make_case_device(mesh, device, seed=1234, dt=dt,
                                      tracers=("T", "S", "T", "S"))
    return _BASE[key]


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
    case = dict(uv=base["uv"] * fac, w_e=base["w_e"] * fac, helem=base["helem"],
                hnode=base["hnode"], hnode_new=base["hnode_new"], T=Ts,
                T_old=[t.clone() for t in Ts], dt=dt)
    if spec.adjoint:
        case["target"] = [t + noise(0.2, t) for t in Ts]
    return case


def adjoint_of_rollout(a, dt, to_host=True):
    """Gradients of sum_surface (T_final - T_target)^2 with respect to the initial
    tracers (T and T_old of step 0) and uv, through the stored trajectory."""
    import torch
    from .adjoint import advect_fct_vjp
    inst, mesh = a.inst, a.mesh
    surf = torch.from_numpy(mesh.node_layer_mask[:, 0]).to(inst.uv.device)
    fin = a.traj[-1][0]
    g = []
    for Tf, tg in zip(fin, a.case_extra["target"]):
        gj = torch.zeros_like(Tf)
        gj[:, 0] = torch.where(surf, 2.0 * (Tf[:, 0] - tg[:, 0]), torch.zeros_like(Tf[:, 0]))
        g.append(gj)
    gTo = [torch.zeros_like(x) for x in g]
    g_uv = torch.zeros_like(inst.uv)
    for s in range(len(a.traj) - 2, -1, -1):
        T, To = a.traj[s]
        gT, gTo_s, guv, _ = advect_fct_vjp(mesh, inst.uv, inst.w_e, inst.helem, inst.hnode,
                                           inst.hnode_new, T, To, g, dt=dt, need_w=False)
        g_uv += guv
        g = [x + y for x, y in zip(gT, gTo)]
        gTo = gTo_s
    if not to_host:
        return {"g_T0": g, "g_T_old0": gTo, "g_uv": g_uv}
    return {"g_T0": [x.cpu().numpy() for x in g], "g_T_old0": [x.cpu().numpy() for x in gTo],
            "g_uv": g_uv.cpu().numpy()}


def fifo_loop(specs, device, *, meshes=None, dt=2700.0):
    """The plain per-request loop: requests in arrival order, one at a time; each
    builds its plan and inputs, runs its steps with the plain entry point, downloads
    its result and synchronises before the next one starts."""
    import torch
    from .batch import Instance
    from .tracer_adv import FctPlan
    meshes = meshes or MeshCache()
    lat, setup, results = {}, {}, {}
    plans = {}
    t0 = time.perf_counter()
    busy = 0.0
    for spec in sorted(specs, key=lambda s: s.arrival):
        now = time.perf_counter() - t0
        if spec.arrival > now:
            time.sleep(spec.arrival - now)
        tb = time.perf_counter()
        mesh = meshes.get(spec.level, spec.nl)
        case = request_case(mesh, spec, device, dt)
        key = (spec.level, spec.nl, spec.K)
        if key not in plans:
            plans[key] = FctPlan(mesh, device, max_tracers=spec.K)
        inst = Instance.from_case(mesh, case, device, plan=plans[key])
        torch.cuda.synchronize(device)
        setup[spec.rid] = time.perf_counter() - tb
        a = _Active(spec=spec, inst=inst, mesh=mesh,
                    case_extra={"uv": case["uv"], "target": case.get("target")},
                    remaining=spec.nsteps, bytes=0, t_admit=0.0, setup_s=setup[spec.rid])
        if spec.adjoint:
            a.traj.append(([x.clone() for x in inst.T()], [x.clone() for x in inst.T_old()]))
        for _ in range(spec.nsteps):
            inst.step_alone()
            if spec.adjoint:
                a.traj.append(([x.clone() for x in inst.T()], [x.clone() for x in inst.T_old()]))
        res = {"T": [x.cpu().numpy() for x in inst.T()]}
        if spec.adjoint:
            res.update(adjoint_of_rollout(a, dt))
        torch.cuda.synchronize(device)
        results[spec.rid] = res
        lat[spec.rid] = time.perf_counter() - t0 - spec.arrival
        busy += time.perf_counter() - tb
        del inst, a, case
    wall = time.perf_counter() - t0
    for p in plans.values():
        p.close()
    out = summarize(specs, lat, setup, wall, busy, meshes)
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
