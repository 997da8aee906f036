"""Benchmark: FP64 FCT tracer advection on one MI355X.

Workload (default): icosahedral sphere level 9 (10*4**9+2 = 2,621,442 nodes before
~30% land is carved out), 70 levels (nl = 70 interfaces, bottom level 10..70 per
column from a seeded bathymetry), 2 tracers (T, S), dt = 2700 s, seeded
streamfunction velocity (CFL <= 0.5), w by continuity. A "step" is one FCT step of
all tracers; steps are chained (T_old <- T, T <- T_new) with the inputs resident in
HBM. Metric: node-level updates/s (tracers x wet node-layers per step / step time).

Prints one JSON line (rank 0, N = 1 GPU only).
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import platform
import sys
import time
from ctypes import c_void_p
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(level, nl, ntracers, seed, dt, reps=5, numpy_level=5):
    """CPU baselines of the FCT step on this host (the reference's own CPU path is JAX,
    which is not installed, and its sources hold no C to compile):

    * the C loop restatement of the step (oracle/fct_loops.c: the original C kernel's
      loops, gcc -O2) in its OpenMP build on every core this process may use, on a
      smaller mesh of the bench's recipe: median of ``reps`` warm steps (all tracers);
    * the numpy restatement (oracle/fct_numpy.py, the array-op graph the reference's
      jnp code runs) on icosa-``numpy_level``: median of ``reps`` warm steps.

    Both results equal the GPU's bit for bit on the same inputs (the tests check the
    two restatements against each other and the GPU against the numpy one)."""
    import ctypes
    from fesom_jax_amd.inputs import make_case
    from fesom_jax_amd.meshgen import icosahedral_mesh
    from fesom_jax_amd.perfmodel import updates_per_step
    from oracle import fct_numpy as O
    from oracle import loops as LP
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    os.environ["OMP_NUM_THREADS"] = str(cores)
    kinds = ("T", "S", "T", "S")[:ntracers]

    def median_rate(mesh, case, fn):
        def step():
            for j in range(ntracers):
                fn(mesh, case["uv"], case["w_e"], case["helem"], case["hnode"],
                   case["hnode_new"], case["T"][j], case["T_old"][j], dt=dt)
        step()                                       # warm-up
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            step()
            ts.append(time.perf_counter() - t)
        med = float(np.median(ts))
        return updates_per_step(mesh, ntracers) / med, med

    mesh = icosahedral_mesh(level, nl, seed=seed)
    case = make_case(mesh, seed=seed + 1, dt=dt, tracers=kinds)
    omp = lambda *a, **k: LP.advect_one_fct(*a, openmp=True, **k)
    rate, med = median_rate(mesh, case, omp)
    try:
        threads = int(LP._lib(True).omp_get_max_threads())
    except (AttributeError, OSError):
        threads = cores
    nmesh = icosahedral_mesh(numpy_level, nl, seed=seed)
    ncase = make_case(nmesh, seed=seed + 1, dt=dt, tracers=kinds)
    nrate, nmed = median_rate(nmesh, ncase, lambda *a, **k: O.advect_one_fct(*a, **k)[0])
    return {"value": rate, "unit": "node-level updates/s", "cores": threads, "kind": "port",
            "sample": (f"C loop restatement of the reference FCT step (oracle/fct_loops.c, "
                       f"gcc -O2 -fopenmp, {threads} threads), median of {reps} warm steps x "
                       f"{ntracers} tracers on icosa-{level} ({mesh.nod2D} nodes, nl={nl}, "
                       f"{updates_per_step(mesh, ntracers)} updates per step; {med:.3f} s per "
                       f"step) on {cpu_model()}"),
            "numpy": {"value": nrate, "unit": "node-level updates/s", "cores": 1,
                      "sample": (f"numpy restatement (oracle/fct_numpy.py), median of {reps} "
                                 f"warm steps on icosa-{numpy_level} ({nmesh.nod2D} nodes, "
                                 f"nl={nl}; {nmed:.3f} s per step)")}}

def measured_traffic(args, mesh, K):
    """HBM-side bytes per step of the default workload from the committed rocprofv3
    summary (FETCH_SIZE + WRITE_SIZE of the step's kernels, each scaled by the counter's
    reading for a copy kernel of known size); None for other workloads."""
    if (args.level, mesh.nl, K, args.seed) != (9, 70, 2, 2024):
        return None
    p = ROOT / "profiles" / "fct_step_icosa9_pmc.md"
    try:
        for line in p.read_text().splitlines():
            if line.startswith("traffic_bytes_per_step:"):
                return int(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--level", type=int, default=9)
    ap.add_argument("--nl", type=int, default=70)
    ap.add_argument("--tracers", type=int, default=2)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--dt", type=float, default=2700.0)
    ap.add_argument("--cpu-level", type=int, default=7)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-limiter-line", action="store_true",
                    help="plain: skip the second timing on a limiter-binding input")
    ap.add_argument("--workload",
                    choices=("plain", "batched", "serve", "adjoint", "vdiff", "compute_w", "ice"),
                    default="plain",
                    help="plain: one icosa-9 instance (the metric); batched: 64 uneven "
                         "instances in one batched launch vs 64 sequential plain launches; "
                         "vdiff: implicit vertical diffusion of T and S on the plain mesh; "
                         "compute_w: vertical velocity from continuity on the plain mesh; "
                         "ice: sea-ice FCT of the three ice tracers on the plain mesh")
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--shuffle", type=int, default=None, metavar="SEED",
                    help="plain: renumber nodes, triangles and edges at random (the plan's "
                         "locality order must make up for a numbering without locality)")
    ap.add_argument("--fd-directions", type=int, default=20,
                    help="adjoint workload: finite-difference directions checked (0 skips)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed steps, run a 1 GiB read + 1 GiB write elementwise "
                         "kernel (the reference for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--rate", type=float, default=20.0, help="serve: requests per second")
    ap.add_argument("--duration", type=float, default=10.0, help="serve: seconds of arrivals")
    ap.add_argument("--levels", type=str, default="4,7", help="serve: icosa level range")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(instances and meshes would shard across GPUs by process). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from fesom_jax_amd import _lib
    from fesom_jax_amd.inputs import make_case_device
    from fesom_jax_amd.meshgen import icosahedral_mesh
    from fesom_jax_amd.perfmodel import HBM_PEAK_GBS, algorithmic_bytes, updates_per_step
    from fesom_jax_amd.tracer_adv import get_plan

    if not torch.cuda.is_available():
        raise RuntimeError("bench.py needs a ROCm GPU")
    dev = torch.device("cuda", 0)
    if args.workload == "batched":
        return run_batched(args, dev)
    if args.workload == "serve":
        return run_serve(args, dev)
    if args.workload == "adjoint":
        return run_adjoint(args, dev)
    if args.workload == "vdiff":
        return run_vdiff(args, dev)
    if args.workload == "compute_w":
        return run_compute_w(args, dev)
    if args.workload == "ice":
        return run_ice(args, dev)
    K = args.tracers

    t0 = time.time()
    mesh = icosahedral_mesh(args.level, args.nl, seed=args.seed)
    if args.shuffle is not None:
        from fesom_jax_amd.meshgen import shuffled
        mesh = shuffled(mesh, args.shuffle)
    log(f"mesh icosa-{args.level}{' (shuffled numbering)' if args.shuffle is not None else ''}: "
        f"N={mesh.nod2D} E={mesh.elem2D} D={mesh.edge2D} nl={mesh.nl} ({time.time() - t0:.1f}s)")
    t0 = time.time()
    kinds = ("T", "S", "T", "S")[:K]
    case = make_case_device(mesh, dev, seed=args.seed + 1, dt=args.dt, tracers=kinds)
    torch.cuda.synchronize()
    log(f"inputs ({time.time() - t0:.1f}s) cfl={case['cfl']} vel_scale={case['vel_scale']:.4f}")

    uv, w_e, helem = case["uv"], case["w_e"], case["helem"]
    hnode, hnn = case["hnode"], case["hnode_new"]
    # three rotating buffers per tracer: (T_old, T, T_new)
    bufs = [[case["T_old"][j], case["T"][j],
             torch.empty((mesh.nod2D, mesh.nl), dtype=torch.float64, device=dev)]
            for j in range(K)]
    del case
    plan = get_plan(mesh, dev)
    lib = plan.lib
    log(f"plan workspace {plan.workspace_bytes / 2**30:.2f} GiB")

    def make_args(rot):
        olds = (c_void_p * K)(*[bufs[j][rot % 3].data_ptr() for j in range(K)])
        curs = (c_void_p * K)(*[bufs[j][(rot + 1) % 3].data_ptr() for j in range(K)])
        news = (c_void_p * K)(*[bufs[j][(rot + 2) % 3].data_ptr() for j in range(K)])
        a = _lib.FctStepArgs(ntracers=K, dt=args.dt, uv=uv.data_ptr(), w_e=w_e.data_ptr(),
                             helem=helem.data_ptr(), hnode=hnode.data_ptr(),
                             hnode_new=hnn.data_ptr(),
                             T=ctypes.cast(curs, ctypes.POINTER(c_void_p)),
                             T_old=ctypes.cast(olds, ctypes.POINTER(c_void_p)),
                             T_new=ctypes.cast(news, ctypes.POINTER(c_void_p)),
                             w_i=None, Z3d=None, zbar3=None)
        return a, (olds, curs, news)

    rots = [make_args(r) for r in range(3)]
    stream = torch.cuda.current_stream(dev)
    sp = c_void_p(stream.cuda_stream)

    def step(i):
        a = rots[i % 3][0]
        _lib.check(lib.fct_forward(plan.ptr, ctypes.byref(a), sp), "fct_forward")

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    log(f"warmup {args.warmup} steps done")
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    start.record(stream)
    for i in range(args.steps):
        step(args.warmup + i)
    end.record(stream)
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall
    ms = start.elapsed_time(end) / args.steps
    log(f"timed {args.steps} steps: {ms:.3f} ms/step (wall {1e3 * t_wall / args.steps:.3f})")

    if args.copy_calibration:
        x = torch.ones(1 << 27, dtype=torch.float64, device=dev)     # 1 GiB
        y = torch.empty_like(x)
        for _ in range(3):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        log("copy calibration: 3 x (1 GiB read + 1 GiB write) done")
        del x, y
    # sanity: the chained field stays finite and bounded
    last = bufs[0][(args.warmup + args.steps + 1) % 3]
    m = torch.from_numpy(mesh.node_layer_mask).to(dev)
    if not bool(torch.isfinite(last[m]).all()):
        raise RuntimeError("non-finite tracer after the timed steps")

    U = updates_per_step(mesh, K)
    B = algorithmic_bytes(mesh, K)
    value = U / (ms * 1e-3)
    achieved = B / (ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": measured_traffic(args, mesh, K),
                "traffic_source": ("not measured by this run: FETCH_SIZE + WRITE_SIZE from "
                                   "rocprofv3 --pmc runs of this workload, copied from "
                                   "profiles/fct_step_icosa9_pmc.md"),
                "kernel": "fct_step (all launches of one step)",
                "bytes_per_step": B}
    lim = None
    if not args.no_limiter_line:
        lim = limiter_binding(mesh, dict(uv=uv, w_e=w_e, helem=helem, hnode=hnode,
                                         hnode_new=hnn), K, args, dev, B, U)
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(args.cpu_level, args.nl, K, args.seed, args.dt)
        log(f"cpu baseline {cpu['value']:.3e} updates/s")
    out = {
        "metric": "FP64 FCT node-level updates/s + %HBM roofline; served p50/p99 vs plain loop",
        "value": value, "unit": "node-level updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"icosa-{args.level} x nl={mesh.nl}, {K} tracers, plain FCT step"
                               + (", shuffled numbering" if args.shuffle is not None else ""),
                   "nodes": mesh.nod2D, "elements": mesh.elem2D, "edges": mesh.edge2D,
                   "levels": mesh.nl, "tracers": K, "updates_per_step": U, "dt": args.dt,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "limiter_binding": lim,
    }
    print(json.dumps(out), flush=True)


def limiter_binding(mesh, flow, K, args, dev, B, U):
    """The same step timed on an input where the Zalesak limiter binds: the reference's
    limiter-active recipe (a sharp tracer 10 + 5 sin(0.7 n) cos(1.3 k), T_old = 10,
    fesom_jax/tests/test_tracers.py:578-588) with the flow scaled by 1.5. Also reports
    the share of wet edge-levels whose antidiffusive flux the limiter reduces and the
    share of wet node interfaces whose vertical one it reduces (from the single-stage
    entry points on the same input)."""
    import torch
    from fesom_jax_amd import _lib
    from fesom_jax_amd import tracer_adv as S
    from fesom_jax_amd.tracer_adv import get_plan
    wet = torch.from_numpy(mesh.node_layer_mask).to(dev)
    n = torch.arange(mesh.nod2D, device=dev, dtype=torch.float64)[:, None]
    k = torch.arange(mesh.nl, device=dev, dtype=torch.float64)[None, :]
    zero = torch.zeros((), dtype=torch.float64, device=dev)
    uv, w_e = flow["uv"] * 1.5, flow["w_e"] * 1.5
    Ts = [torch.where(wet, 10.0 + 5.0 * torch.sin(n * (0.7 + 0.1 * j)) * torch.cos(k * 1.3), zero)
          for j in range(K)]
    To = [torch.where(wet, 10.0, 0.0).to(torch.float64) for _ in range(K)]
    Tn = [torch.empty_like(Ts[0]) for _ in range(K)]
    plan = get_plan(mesh, dev)
    arr = lambda ts: ctypes.cast((c_void_p * K)(*[t.data_ptr() for t in ts]),
                                 ctypes.POINTER(c_void_p))
    keep = [arr(Ts), arr(To), arr(Tn)]
    a = _lib.FctStepArgs(ntracers=K, dt=args.dt, uv=uv.data_ptr(), w_e=w_e.data_ptr(),
                         helem=flow["helem"].data_ptr(), hnode=flow["hnode"].data_ptr(),
                         hnode_new=flow["hnode_new"].data_ptr(), T=keep[0], T_old=keep[1],
                         T_new=keep[2], w_i=None, Z3d=None, zbar3=None)
    stream = torch.cuda.current_stream(dev)
    sp = c_void_p(stream.cuda_stream)
    for _ in range(args.warmup):
        _lib.check(plan.lib.fct_forward(plan.ptr, ctypes.byref(a), sp), "fct_forward")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        _lib.check(plan.lib.fct_forward(plan.ptr, ctypes.byref(a), sp), "fct_forward")
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    # binding shares from the stages of the first tracer's step
    T, Tol, dt = Ts[0], To[0], args.dt
    fh = S.adv_flux_hor(mesh, uv, flow["helem"], T)
    fv = S.adv_flux_ver(mesh, w_e, T)
    LO = S.compute_fct_lo(mesh, fh, fv, T, flow["hnode"], flow["hnode_new"], dt=dt)
    ttf = S.tracer_ab(T, Tol)[0]
    eud = S.fill_up_dn_grad(mesh, S.tracer_gradient_elements(mesh, T))
    adf_h = S.adv_flux_hor_ho(mesh, uv, flow["helem"], ttf, eud) - fh
    del eud
    adf_v = S.adv_flux_ver_ho(mesh, w_e, ttf) - fv
    lim_h, lim_v = S.zalesak_limit(mesh, T, LO, adf_h, adf_v, flow["hnode_new"], dt=dt)
    ewet = adf_h != 0.0
    vwet = torch.from_numpy(mesh.node_iface_mask).to(dev) & (adf_v != 0.0)
    share_h = float((ewet & (lim_h.abs() < adf_h.abs())).sum()) / max(1, int(ewet.sum()))
    share_v = float((vwet & (lim_v.abs() < adf_v.abs())).sum()) / max(1, int(vwet.sum()))
    del fh, fv, LO, ttf, adf_h, adf_v, lim_h, lim_v
    log(f"limiter-binding input: {ms:.3f} ms/step, limiter reduces {100 * share_h:.1f} % of "
        f"edge fluxes and {100 * share_v:.1f} % of vertical fluxes")
    achieved = B / (ms * 1e-3) / 1e9
    return {"input": "sharp tracer 10 + 5 sin(0.7 n) cos(1.3 k), T_old = 10, flow x 1.5",
            "ms_per_step": ms, "value": U / (ms * 1e-3),
            "frac": round(achieved / 8000.0, 5),
            "limited_share_edge_fluxes": share_h, "limited_share_vertical_fluxes": share_v}


def run_batched(args, dev):
    """64 independent instances: icosahedral levels 5-8 (equal odds, i.e. log-uniform
    over the four node counts), nl 47 or 70, 2-4 tracers, per-instance seeded fields.
    Instances with the same (level, nl) share the generated mesh geometry but own their
    plan, workspace, flow and tracers. Times S steps of (a) every instance through its
    own plain fct_forward launches, one after the other, and (b) one batched launch set
    per step replayed as a HIP graph, with the launch shape picked by the autotuner."""
    import torch
    from fesom_jax_amd import autotune
    from fesom_jax_amd.batch import FctBatch, Instance
    from fesom_jax_amd.inputs import make_case_device
    from fesom_jax_amd.meshgen import icosahedral_mesh
    from fesom_jax_amd.perfmodel import HBM_PEAK_GBS, algorithmic_bytes

    rng = np.random.default_rng(args.seed)
    specs = [(int(rng.integers(5, 9)), int(rng.choice([47, 70])), int(rng.integers(2, 5)))
             for _ in range(args.instances)]
    meshes, insts, B = {}, [], 0
    t0 = time.time()
    for i, (lv, nl, K) in enumerate(specs):
        if (lv, nl) not in meshes:
            meshes[(lv, nl)] = icosahedral_mesh(lv, nl, seed=args.seed + 10 * lv + nl)
        mesh = meshes[(lv, nl)]
        case = make_case_device(mesh, dev, seed=args.seed + 1000 + i, dt=args.dt,
                                tracers=("T", "S", "T", "S")[:K])
        insts.append(Instance.from_case(mesh, case, dev))
        B += algorithmic_bytes(mesh, K)
        del case
    torch.cuda.synchronize()
    U = sum(inst.updates for inst in insts)
    log(f"{len(insts)} instances ({time.time() - t0:.1f}s): levels "
        f"{sorted(set(s[0] for s in specs))}, {U} updates/step, "
        f"device memory {torch.cuda.memory_allocated(dev) / 2**30:.1f} GiB (torch) ")
    stream = torch.cuda.current_stream(dev)

    def seq_step():
        for inst in insts:
            inst.step_alone(stream)

    batch = FctBatch(insts)
    best, table = autotune.tune_batch(
        batch, stream, reps=2, warmup=1, search="descent",
        progress=lambda cfg, ms: log(f"tune {autotune.describe(cfg)}: {ms:.2f} ms"))
    log(f"autotuner: {autotune.describe(best)} "
        f"({min(ms for _, ms in table):.3f} .. {max(ms for _, ms in table):.3f} ms/step)")
    batch.prepare()
    ms_seq = autotune.time_ms(seq_step, stream, reps=args.steps, warmup=args.warmup)
    batch.prepare()
    ms_bat = autotune.time_ms(lambda: batch.step(graph=True, stream=stream), stream,
                              reps=args.steps, warmup=args.warmup)
    log(f"sequential plain {ms_seq:.3f} ms/step, batched graph {ms_bat:.3f} ms/step")
    for inst in insts:
        for t in inst.T():
            if not bool(torch.isfinite(t).all()):
                raise RuntimeError("non-finite tracer after the batched steps")
    achieved = B / (ms_bat * 1e-3) / 1e9
    out = {
        "metric": "FP64 FCT node-level updates/s, 64 uneven instances in one batched "
                  "launch (vs 64 sequential plain launches)",
        "value": U / (ms_bat * 1e-3), "unit": "node-level updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_bat,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{len(insts)} instances, icosa 5-8, nl 47/70, 2-4 tracers",
                   "instances": [{"level": lv, "nl": nl, "tracers": K} for lv, nl, K in specs],
                   "updates_per_step": U, "dt": args.dt, "parallelism": "single GPU",
                   "tuning": autotune.describe(best)},
        "sequential_plain": {"value": U / (ms_seq * 1e-3), "ms_per_step": ms_seq},
        "speedup_vs_sequential": ms_seq / ms_bat,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": None, "kernel": "batched fct step (all launches of one step)",
                     "bytes_per_step": B},
        "cpu_baseline": None,
    }
    batch.close()
    print(json.dumps(out), flush=True)


def run_serve(args, dev):
    """A Poisson stream of forward (70%) and forward+adjoint (30%) requests of 1-24 steps,
    2-4 tracers, nl 47/70, icosa levels in --levels, served with continuous batching
    (200 GB HBM budget) and, separately, by the plain per-request FIFO loop."""
    import torch
    from fesom_jax_amd.serve import MeshCache, Server, fifo_loop, make_stream
    lo, hi = (int(x) for x in args.levels.split(","))
    specs = make_stream(args.rate, args.duration, levels=(lo, hi), seed=args.seed)
    meshes = MeshCache(seed=args.seed)
    t0 = time.time()
    for lv in range(lo, hi + 1):
        for nl in (47, 70):
            meshes.get(lv, nl)
    log(f"{len(specs)} requests over {args.duration}s; meshes ready ({time.time() - t0:.1f}s)")
    # warm-up, untimed: one forward+adjoint request per (level, nl) through both paths,
    # so that meshes and base inputs exist and both code paths have run before either
    # is timed (each path's plans are released before the next one runs)
    from fesom_jax_amd.serve import RequestSpec
    warm = [RequestSpec(rid=i, arrival=0.0, level=lv, nl=nl, K=2, nsteps=2, adjoint=True,
                        seed=i) for i, (lv, nl) in enumerate(
                            (lv, nl) for lv in range(lo, hi + 1) for nl in (47, 70))]
    warm_srv = Server(dev, meshes=meshes)
    warm_srv.run(warm)
    warm_srv.close()
    torch.cuda.empty_cache()
    fifo_loop(warm, dev, meshes=meshes)
    torch.cuda.empty_cache()
    torch.cuda.synchronize()
    log("warm-up done")
    srv = Server(dev, budget_bytes=200 << 30, max_active=64, meshes=meshes)
    served = srv.run(specs, log=log)
    srv.close()
    log(f"served: {json.dumps(served)}")
    torch.cuda.empty_cache()
    fifo = fifo_loop(specs, dev, meshes=meshes)
    fifo.pop("results")
    log(f"fifo: {json.dumps(fifo)}")
    out = {
        "metric": "FP64 FCT served p50/p99 latency and node-level updates/s vs plain FIFO loop",
        "value": served["updates_per_s"], "unit": "node-level updates/s", "n_gpus": 1,
        "steps": None, "warmup": None, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"Poisson {args.rate}/s for {args.duration}s, icosa {lo}-{hi}, "
                               "nl 47/70, 2-4 tracers, 1-24 steps, 30% with adjoint",
                   "requests": len(specs), "hbm_budget_GB": 200, "parallelism": "single GPU"},
        "served": served, "fifo_plain_loop": fifo,
        "speedup_updates_per_s": served["updates_per_s"] / fifo["updates_per_s"],
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def run_adjoint(args, dev):
    """12 chained steps on icosa-7 x nl=47 with 2 tracers, loss = sum over wet surface
    nodes of (T_final - T_target)^2 (T_target = T0 + 0.2 N(0,1)), gradients with respect
    to the initial T, S and uv through the hand-written adjoint (per-step recompute).
    Times forward and forward+backward, and checks the gradient against central
    differences of the GPU forward along 20 seeded directions (h = 1e-6 relative)."""
    import torch
    from fesom_jax_amd.adjoint import rollout
    from fesom_jax_amd.inputs import make_case_device
    from fesom_jax_amd.meshgen import icosahedral_mesh
    from fesom_jax_amd.perfmodel import updates_per_step
    level, nl, K, nsteps = 7, 47, 2, 12
    mesh = icosahedral_mesh(level, nl, seed=args.seed)
    case = make_case_device(mesh, dev, seed=args.seed + 1, dt=args.dt, tracers=("T", "S"))
    g = torch.Generator(device=dev).manual_seed(args.seed)
    f64 = dict(dtype=torch.float64, device=dev)
    nlm = torch.from_numpy(mesh.node_layer_mask).to(dev)
    zero = torch.zeros((), **f64)
    target = [t + torch.where(nlm, 0.2 * torch.randn(t.shape, generator=g, **f64), zero)
              for t in case["T"]]
    surf = nlm[:, 0].to(torch.float64)
    fixed = (case["w_e"], case["helem"], case["hnode"], case["hnode_new"])

    def loss_of(T0, To0, uv):
        fin = rollout(mesh, uv, *fixed, T0, To0, nsteps, dt=args.dt)
        return sum((((f[:, 0] - tg[:, 0]) ** 2) * surf).sum() for f, tg in zip(fin, target))

    T0 = [t.clone().requires_grad_(True) for t in case["T"]]
    To0 = [t.clone() for t in case["T"]]
    uv = case["uv"].clone().requires_grad_(True)
    # warm-up (untimed): the first backward allocates the plan's adjoint workspace
    torch.autograd.grad(loss_of(T0, To0, uv), [*T0, uv])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.no_grad():
        loss_of([t.detach() for t in T0], To0, uv.detach())
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    L0 = loss_of(T0, To0, uv)
    grads = torch.autograd.grad(L0, [*T0, uv])
    L0 = float(L0.detach())
    torch.cuda.synchronize()
    t_fb = time.perf_counter() - t0
    log(f"forward {t_fwd:.3f}s, forward+adjoint {t_fb:.3f}s, loss {float(L0):.6e}")
    # one step's backward alone (fct_backward, both tracers), HIP events on the stream
    # the library launches on, against its compulsory bytes
    from fesom_jax_amd.adjoint import advect_fct_vjp
    from fesom_jax_amd.perfmodel import HBM_PEAK_GBS, adjoint_bytes
    gnew = [torch.randn(t.shape, generator=g, **f64) for t in case["T"]]
    bwd = lambda: advect_fct_vjp(mesh, case["uv"], *fixed, case["T"], To0, gnew, dt=args.dt)
    bwd()
    reps = 10
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        bwd()
    e1.record()
    torch.cuda.synchronize()
    t_bwd = e0.elapsed_time(e1) / reps * 1e-3
    B_adj = adjoint_bytes(mesh, K)
    log(f"one step's backward {1e3 * t_bwd:.3f} ms, {B_adj / t_bwd / 1e9:.1f} GB/s compulsory")
    # central differences along 20 seeded directions, each on one random column (a wet
    # node's T or S column, or a triangle's uv column) near the surface signal, so that
    # few limiter switches lie on the segment; step h = 1e-6 * max|x| (and 1e-5, 1e-7:
    # the best agreement of the three is kept, which skips a switch inside one stencil)
    rel = []
    rng = np.random.default_rng(args.seed + 7)
    base = [x.detach() for x in T0] + [uv.detach()]
    wet_nodes = np.nonzero(mesh.node_layer_mask[:, 0])[0]
    with torch.no_grad():
        for r in range(args.fd_directions):
            which = r % 3
            d = torch.zeros_like(base[which])
            if which < 2:
                nd = int(rng.choice(wet_nodes))
                col = rng.standard_normal(mesh.nl) * mesh.node_layer_mask[nd]
                d[nd] = torch.from_numpy(col).to(dev)
            else:
                el = int(rng.integers(mesh.elem2D))
                col = rng.standard_normal((mesh.nl, 2)) * mesh.elem_layer_mask[el][:, None]
                d[el] = torch.from_numpy(col).to(dev)
            ad = float((grads[which] * d).sum())
            best = np.inf
            for hr in (1e-6, 1e-5, 1e-7):
                h = hr * float(base[which].abs().max())
                def at(sign):
                    xs = [b.clone() for b in base]
                    xs[which] = xs[which] + sign * h * d
                    return float(loss_of(xs[:2], To0, xs[2]))
                fd = (at(1.0) - at(-1.0)) / (2 * h)
                best = min(best, abs(fd - ad) / max(abs(fd), abs(ad), 1e-300))
            rel.append(best)
    rel = np.array(rel if rel else [np.nan])
    log(f"FD check: median rel {np.median(rel):.2e}, max {rel.max():.2e}, "
        f"{int((rel < 1e-5).sum())}/{args.fd_directions} below 1e-5")
    U = updates_per_step(mesh, K) * nsteps
    out = {
        "metric": "FP64 FCT forward+adjoint rollout (12 steps) node-level updates/s",
        "value": U / t_fb, "unit": "node-level updates/s (forward updates / forward+backward time)",
        "n_gpus": 1, "steps": nsteps, "warmup": 0, "ms_per_step": 1e3 * t_fb / nsteps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"icosa-{level} x nl={nl}, {K} tracers, {nsteps} steps + adjoint",
                   "nodes": mesh.nod2D, "updates_forward": U},
        "forward_s": t_fwd, "forward_backward_s": t_fb, "backward_over_forward": (t_fb - t_fwd) / t_fwd,
        "fd_check": {"directions": args.fd_directions,
                     "kind": "one random column each, h in 1e-6/1e-5/1e-7",
                     "median_rel": float(np.median(rel)),
                     "max_rel": float(rel.max()), "below_1e-5": int((rel < 1e-5).sum())},
        "backward_step_ms": 1e3 * t_bwd,
        "roofline": {"bound": "hbm", "achieved": round(B_adj / t_bwd / 1e9, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(B_adj / t_bwd / 1e9 / HBM_PEAK_GBS, 5), "traffic": None,
                     "kernel": "fct_backward (all launches of one step's backward, 2 tracers)",
                     "bytes_per_step": B_adj},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def run_vdiff(args, dev):
    """Implicit vertical diffusion (the step after the FCT) of T and S with one Kv on the
    plain workload's mesh, in place, repeated. Algorithmic bytes per call: T and S read
    and written in full (2 x 2 x 8 B per node-level: dry levels are copied), Kv, area,
    areasvol and hnode_new read on the wet layers plus the interface below
    (4 x 8 B per such entry)."""
    import torch
    from fesom_jax_amd.inputs import make_case_device
    from fesom_jax_amd.meshgen import icosahedral_mesh
    from fesom_jax_amd.perfmodel import HBM_PEAK_GBS
    from fesom_jax_amd.tracer_diff import diffuse_tracers
    K = 2
    mesh = icosahedral_mesh(args.level, args.nl, seed=args.seed)
    case = make_case_device(mesh, dev, seed=args.seed + 1, dt=args.dt, tracers=("T", "S"))
    g = torch.Generator(device=dev).manual_seed(args.seed + 3)
    iface = torch.from_numpy(mesh.node_iface_mask).to(dev)
    Kv = torch.where(iface, 1e-5 + 3e-2 * torch.rand((mesh.nod2D, mesh.nl), generator=g,
                                                    dtype=torch.float64, device=dev),
                     torch.zeros((), dtype=torch.float64, device=dev))
    Ts = [case["T"][0].clone(), case["T"][1].clone()]
    hnn = case["hnode_new"]
    del case
    wet = mesh.node_layer_mask
    n_wet = int(wet.sum())
    cols_wet = int(wet.any(axis=1).sum())
    U = K * n_wet
    B = 8 * (2 * K * mesh.nod2D * mesh.nl + 4 * (n_wet + cols_wet))
    stream = torch.cuda.current_stream(dev)

    def call():
        diffuse_tracers(mesh, Ts, Kv, hnn, dt=args.dt, floor_tracer=1, out=Ts)

    for _ in range(max(1, args.warmup)):
        call()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(args.steps):
        call()
    end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / args.steps
    m = torch.from_numpy(wet).to(dev)
    if not bool(torch.isfinite(Ts[0][m]).all()):
        raise RuntimeError("non-finite tracer after the timed calls")
    achieved = B / (ms * 1e-3) / 1e9
    log(f"vdiff icosa-{args.level}: {ms:.3f} ms/call, {achieved:.0f} GB/s algorithmic")
    cpu = None
    if not args.no_cpu_baseline:
        from fesom_jax_amd.inputs import make_case
        from oracle import vert_diff as VD
        cm = icosahedral_mesh(args.cpu_level, args.nl, seed=args.seed)
        cc = make_case(cm, seed=args.seed + 1, dt=args.dt)
        rng = np.random.default_rng(args.seed + 3)
        ckv = np.where(cm.node_iface_mask, 1e-5 + 3e-2 * rng.random((cm.nod2D, cm.nl)), 0.0)
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 10.0 or reps < 1:
            VD.impl_vert_diff(cm, cc["T"][0], cc["T"][1], ckv, cc["hnode_new"], dt=args.dt)
            reps += 1
        sec = (time.perf_counter() - t0) / reps
        cu = 2 * int(cm.node_layer_mask.sum())
        cpu = {"value": cu / sec, "unit": "node-level updates/s", "cores": 1, "kind": "port",
               "sample": f"numpy oracle impl_vert_diff, icosa-{args.cpu_level} x nl={cm.nl}, "
                         f"{reps} calls in {sec * reps:.1f} s on {cpu_model()}"}
    out = {
        "metric": "FP64 implicit vertical diffusion node-level updates/s",
        "value": U / (ms * 1e-3), "unit": "node-level updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"icosa-{args.level} x nl={mesh.nl}, T and S, impl_vert_diff "
                               f"+ salinity floor, in place", "nodes": mesh.nod2D,
                   "levels": mesh.nl, "updates_per_call": U, "dt": args.dt},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": None, "kernel": "k_vdiff", "bytes_per_call": B},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def run_compute_w(args, dev):
    """Vertical velocity by continuity (the w the step takes) on the plain workload's
    mesh. Algorithmic bytes per call: uv and helem on wet element layers (24 B each),
    area on wet node interfaces (8 B), w written in full (8 B per node-level), and the
    edge tables (edge triangles, cross vectors, CSR entries: 56 B per edge)."""
    import torch
    from fesom_jax_amd.ale import compute_w
    from fesom_jax_amd.inputs import make_case_device
    from fesom_jax_amd.meshgen import icosahedral_mesh
    from fesom_jax_amd.perfmodel import HBM_PEAK_GBS
    mesh = icosahedral_mesh(args.level, args.nl, seed=args.seed)
    case = make_case_device(mesh, dev, seed=args.seed + 1, dt=args.dt, tracers=("T",))
    uv, helem = case["uv"], case["helem"]
    del case
    n_le = int(mesh.elem_layer_mask.sum())
    n_in = int(mesh.node_iface_mask.sum())
    B = 24 * n_le + 8 * n_in + 8 * mesh.nod2D * mesh.nl + 56 * mesh.edge2D
    stream = torch.cuda.current_stream(dev)
    for _ in range(max(1, args.warmup)):
        w = compute_w(mesh, uv, helem)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(args.steps):
        w = compute_w(mesh, uv, helem)
    end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / args.steps
    if not bool(torch.isfinite(w).all()):
        raise RuntimeError("non-finite w")
    achieved = B / (ms * 1e-3) / 1e9
    log(f"compute_w icosa-{args.level}: {ms:.3f} ms/call, {achieved:.0f} GB/s algorithmic")
    out = {
        "metric": "FP64 compute_w node-interface values/s",
        "value": n_in / (ms * 1e-3), "unit": "node interfaces/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"icosa-{args.level} x nl={mesh.nl}, compute_w",
                   "nodes": mesh.nod2D, "levels": mesh.nl, "interfaces": n_in},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": None, "kernel": "kf_vflux + k_compute_w_vf", "bytes_per_call": B},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def run_ice(args, dev):
    """Sea-ice FCT of a_ice, m_ice, m_snow on the plain workload's mesh (surface nodes),
    ice_dt = 4000 s, velocity up to 0.5 m/s. Algorithmic bytes per call: the five input
    and three output fields (64 B per node), node area and level range (16 B per node),
    the element tables (vertices 12 B, gradients 48 B, area and its sqrt factor 16 B,
    level range 8 B, node->element entries 12 B: 96 B per element)."""
    import torch
    from fesom_jax_amd.ice_adv import IceConfig, fct_solve
    from fesom_jax_amd.inputs import make_ice_case
    from fesom_jax_amd.meshgen import icosahedral_mesh
    from fesom_jax_amd.perfmodel import HBM_PEAK_GBS
    mesh = icosahedral_mesh(args.level, args.nl, seed=args.seed)
    c = make_ice_case(mesh, seed=args.seed + 5, speed=0.5)
    f = {k: torch.from_numpy(v).to(dev) for k, v in c.items()}
    cfg = IceConfig(ice_dt=4000.0)
    N, E = mesh.nod2D, mesh.elem2D
    B = 80 * N + 96 * E
    stream = torch.cuda.current_stream(dev)

    def call():
        return fct_solve(cfg, mesh, f["a_ice"], f["m_ice"], f["m_snow"], f["u_ice"], f["v_ice"])

    for _ in range(max(1, args.warmup)):
        out = call()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(args.steps):
        out = call()
    end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / args.steps
    if not all(bool(torch.isfinite(o).all()) for o in out):
        raise RuntimeError("non-finite ice tracer")
    achieved = B / (ms * 1e-3) / 1e9
    log(f"ice FCT icosa-{args.level}: {ms:.3f} ms/call, {achieved:.0f} GB/s algorithmic")
    result = {
        "metric": "FP64 sea-ice FCT node updates/s (3 tracers per node)",
        "value": 3 * N / (ms * 1e-3), "unit": "node-tracer updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"icosa-{args.level} surface, sea-ice FCT, ice_dt 4000 s",
                   "nodes": N, "elements": E},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": None, "kernel": "k_ice_* (six launches: the packing of the inputs, then five passes)", "bytes_per_call": B},
        "cpu_baseline": None,
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
