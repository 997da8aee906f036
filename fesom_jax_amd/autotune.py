"""Autotuner of the fused path's launch shapes.

Knobs (``FCT_TUNE_*`` in ``include/fesom_fct.h``), one axis each: the XCD remap of
workgroups and its range length (which XCD's L2 sees which part of the mesh), the
wavefronts per workgroup, the edge tiles of the two edge passes (edges per wavefront),
and, for the adf_h pass of a single plan, the level block of its edge-patch kernel
(levels staged per chunk; 0 = one wavefront per edge column) and the threads per edge.
The node passes' level block is fixed by the hardware: one 64-lane wavefront per 64
levels. Results never depend on the knobs; only time does. The tuner times
configurations on the caller's own workload with HIP events on the launch stream:
every one of a small space, or (``search="descent"``) one axis at a time from the
library's defaults, keeping each axis's fastest value, until a sweep changes nothing.
No reference counterpart (XLA picks its own shapes).
"""

from __future__ import annotations

import itertools
import json
import os

from . import _lib

SPACE = {
    _lib.FCT_TUNE_XCD: (1, 8),
    _lib.FCT_TUNE_XCD_CHUNK: (0, 8, 32),
    _lib.FCT_TUNE_WAVES_PER_GROUP: (1, 2, 4),
    _lib.FCT_TUNE_VFLUX_EDGES: (2, 4),
    _lib.FCT_TUNE_ADFH_EDGES: (1, 2),
    _lib.FCT_TUNE_ADFH_CHUNK: (0, 4, 8),
    _lib.FCT_TUNE_ADFH_THREADS: (1, 2),
}
# the axes a batch uses (its adf_h pass runs one wavefront per edge column)
BATCH_SPACE = {k: v for k, v in SPACE.items()
               if k not in (_lib.FCT_TUNE_ADFH_CHUNK, _lib.FCT_TUNE_ADFH_THREADS)}
DEFAULT = {_lib.FCT_TUNE_XCD: 8, _lib.FCT_TUNE_XCD_CHUNK: 8, _lib.FCT_TUNE_WAVES_PER_GROUP: 4,
           _lib.FCT_TUNE_VFLUX_EDGES: 4, _lib.FCT_TUNE_ADFH_EDGES: 2,
           _lib.FCT_TUNE_ADFH_CHUNK: 8, _lib.FCT_TUNE_ADFH_THREADS: 1}
NAMES = {_lib.FCT_TUNE_XCD: "xcd", _lib.FCT_TUNE_XCD_CHUNK: "xcd_chunk",
         _lib.FCT_TUNE_WAVES_PER_GROUP: "waves_per_group",
         _lib.FCT_TUNE_VFLUX_EDGES: "vflux_edges", _lib.FCT_TUNE_ADFH_EDGES: "adfh_edges",
         _lib.FCT_TUNE_ADFH_CHUNK: "adfh_level_chunk", _lib.FCT_TUNE_ADFH_THREADS: "adfh_threads"}

_CACHE: dict = {}


def time_ms(run, stream, *, reps=3, warmup=1) -> float:
    """Mean milliseconds of ``run()`` over ``reps`` calls, HIP events on ``stream``."""
    import torch
    for _ in range(warmup):
        run()
    stream.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        run()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def tune(set_knob, run, stream, *, space=None, reps=3, warmup=1, key=None, cache_path=None,
         search="grid", max_sweeps=3, progress=None):
    """Find the fastest configuration of ``space`` (default ``SPACE``) and leave it
    applied. ``set_knob(knob, value)`` applies one knob; ``run()`` launches one step on
    ``stream``. ``search="grid"`` times every configuration; ``"descent"`` starts from
    the library's defaults and times one axis at a time, keeping each axis's fastest
    value, for at most ``max_sweeps`` sweeps or until a sweep changes nothing. Returns
    ``(best, table)`` with ``best`` a {knob: value} dict and ``table`` a list of
    ({knob: value}, ms). A ``key`` reuses an earlier result from this process (or from
    the JSON file ``cache_path``). ``progress(cfg, ms)`` is called after each timing."""
    space = SPACE if space is None else space
    if key is not None:
        hit = _lookup(key, cache_path)
        if hit is not None:
            for k, v in hit.items():
                set_knob(k, v)
            return hit, []
    knobs = sorted(space)
    table = []

    def timed(cfg):
        for k, v in cfg.items():
            set_knob(k, v)
        ms = time_ms(run, stream, reps=reps, warmup=warmup)
        table.append((dict(cfg), ms))
        if progress is not None:
            progress(cfg, ms)
        return ms

    if search == "grid":
        for values in itertools.product(*(space[k] for k in knobs)):
            timed(dict(zip(knobs, values)))
        best = min(table, key=lambda r: r[1])[0]
    elif search == "descent":
        best = {k: DEFAULT.get(k, space[k][0]) for k in knobs}
        best_ms = timed(best)
        for _ in range(max_sweeps):
            changed = False
            for k in knobs:
                for v in space[k]:
                    if v == best[k]:
                        continue
                    cfg = dict(best)
                    cfg[k] = v
                    ms = timed(cfg)
                    if ms < best_ms:
                        best, best_ms, changed = cfg, ms, True
            if not changed:
                break
    else:
        raise ValueError(f"unknown search {search!r}")
    for k, v in best.items():
        set_knob(k, v)
    if key is not None:
        _store(key, best, cache_path)
    return best, table


def tune_plan(plan, args, stream, **kw):
    """Tune a plan's single-instance fused launches on fixed step arguments."""
    import ctypes
    from ctypes import c_void_p
    sp = c_void_p(stream.cuda_stream)

    def set_knob(k, v):
        _lib.check(plan.lib.fct_plan_set_tuning(plan.ptr, k, v), "fct_plan_set_tuning")

    def run():
        _lib.check(plan.lib.fct_forward_ex(plan.ptr, ctypes.byref(args), _lib.FCT_PATH_FUSED,
                                           sp), "fct_forward_ex")
    return tune(set_knob, run, stream, **kw)


def tune_batch(batch, stream, *, graph=True, **kw):
    """Tune a batch (the timed runs advance its instances) over ``BATCH_SPACE``."""
    kw.setdefault("space", BATCH_SPACE)
    return tune(batch.set_tuning, lambda: batch.step(graph=graph, stream=stream), stream, **kw)


def describe(cfg) -> dict:
    return {NAMES.get(k, str(k)): v for k, v in cfg.items()}


def _lookup(key, path):
    key = str(key)
    if key in _CACHE:
        return _CACHE[key]
    if path and os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
        if key in data:
            _CACHE[key] = {int(k): int(v) for k, v in data[key].items()}
            return _CACHE[key]
    return None


def _store(key, best, path):
    key = str(key)
    _CACHE[key] = dict(best)
    if path:
        data = {}
        if os.path.exists(path):
            with open(path) as f:
                data = json.load(f)
        data[key] = {str(k): int(v) for k, v in best.items()}
        with open(path, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
