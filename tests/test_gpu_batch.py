"""Batched FCT steps (fct_batch_*) against stepping each instance alone, bit for bit."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _instances(dev, specs, pi_mesh=None, pi_case=None):
    from fesom_jax_amd.batch import Instance
    from fesom_jax_amd.inputs import make_case
    from fesom_jax_amd.meshgen import icosahedral_mesh
    out = []
    for level, nl, K, seed, *deg8 in specs:
        if level == "pi":
            mesh, case = pi_mesh, dict(pi_case)
            case["T"], case["T_old"] = case["T"][:K], case["T_old"][:K]
        else:
            mesh = icosahedral_mesh(level, nl, seed=seed, degree8=deg8[0] if deg8 else 0)
            case = make_case(mesh, seed=seed + 1, dt=2700.0, tracers=("T", "S", "T", "S")[:K])
            # a distinct T_old so that AB2 matters
            case["T_old"] = [0.999 * t for t in case["T"]]
        out.append(Instance.from_case(mesh, case, dev))
    return out


# the last instance has nodes of 8 neighbours (8 table slots) at nl = 70
SPECS = [("pi", 48, 2, 0), (3, 48, 1, 5), (3, 70, 3, 7), (4, 70, 2, 9), (3, 70, 2, 11, 12)]


def _clone_state(insts):
    return [[[b.clone() for b in bufs] for bufs in inst.bufs] for inst in insts]


def _restore(insts, state, phases):
    for inst, bufs, ph in zip(insts, state, phases):
        for dst, src in zip(inst.bufs, bufs):
            for d, s in zip(dst, src):
                d.copy_(s)
        inst.phase = ph


def _snapshot(insts):
    return [[t.cpu().numpy().copy() for t in inst.T()] for inst in insts]


@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_alone(graph, pi_mesh, pi_case, torch_dev):
    import torch
    from fesom_jax_amd.batch import FctBatch
    insts = _instances(torch_dev, SPECS, pi_mesh, pi_case)
    state0, ph0 = _clone_state(insts), [i.phase for i in insts]
    # reference: each instance alone, 4 steps (wraps the 3-slot rotation)
    ref = []
    for _ in range(4):
        for inst in insts:
            inst.step_alone()
        torch.cuda.synchronize()
        ref.append(_snapshot(insts))
    _restore(insts, state0, ph0)
    batch = FctBatch(insts)
    for s in range(4):
        batch.step(graph=graph)
        torch.cuda.synchronize()
        got = _snapshot(insts)
        for i, (g, r) in enumerate(zip(got, ref[s])):
            for j, (a, b) in enumerate(zip(g, r)):
                assert np.array_equal(a, b), f"step {s} instance {i} tracer {j} differs"
    batch.close()


def test_batch_tuning_does_not_change_results(torch_dev):
    import torch
    from fesom_jax_amd import _lib
    from fesom_jax_amd.batch import FctBatch
    insts = _instances(torch_dev, [(3, 48, 2, 11), (3, 70, 1, 13)])
    state0, ph0 = _clone_state(insts), [i.phase for i in insts]
    batch = FctBatch(insts)
    results = []
    for cfg in [{}, {_lib.FCT_TUNE_XCD: 1, _lib.FCT_TUNE_WAVES_PER_GROUP: 1},
                {_lib.FCT_TUNE_VFLUX_EDGES: 2, _lib.FCT_TUNE_ADFH_EDGES: 1,
                 _lib.FCT_TUNE_WAVES_PER_GROUP: 2},
                {_lib.FCT_TUNE_XCD: 8, _lib.FCT_TUNE_XCD_CHUNK: 5}]:
        _restore(insts, state0, ph0)
        for k, v in cfg.items():
            batch.set_tuning(k, v)
        batch.prepare()
        batch.step(graph=True)
        batch.step(graph=True)
        torch.cuda.synchronize()
        results.append(_snapshot(insts))
    for r in results[1:]:
        for g, b in zip(r, results[0]):
            for x, y in zip(g, b):
                assert np.array_equal(x, y)
    batch.close()


def test_graph_replays_with_eager_work_between(torch_dev):
    """Replays of the captured graph interleaved with unrelated eager work and syncs
    give the same states as plain launches."""
    import torch
    from fesom_jax_amd.batch import FctBatch
    insts = _instances(torch_dev, [(3, 48, 2, 21), (3, 48, 2, 23)])
    state0, ph0 = _clone_state(insts), [i.phase for i in insts]
    batch = FctBatch(insts)
    for _ in range(6):
        batch.step(graph=False)
    torch.cuda.synchronize()
    want = _snapshot(insts)
    _restore(insts, state0, ph0)
    batch.prepare()
    junk = torch.randn(1 << 20, device=torch_dev, dtype=torch.float64)
    for _ in range(6):
        batch.step(graph=True)
        junk = junk * 1.0001 + junk.sum() * 0.0   # eager reductions between replays
        torch.cuda.synchronize()
    got = _snapshot(insts)
    for g, w in zip(got, want):
        for x, y in zip(g, w):
            assert np.array_equal(x, y)
    batch.close()


def test_batch_rejects_bad_input(pi_mesh, pi_case, torch_dev):
    from fesom_jax_amd import _lib
    from fesom_jax_amd.batch import FctBatch
    insts = _instances(torch_dev, [("pi", 48, 2, 0)], pi_mesh, pi_case)
    with pytest.raises(ValueError):
        FctBatch([insts[0], insts[0]])            # one workspace, two instances
    batch = FctBatch(insts)
    with pytest.raises(ValueError):
        batch.set_tuning(_lib.FCT_TUNE_WAVES_PER_GROUP, 3)
    batch.close()


def test_autotuner_picks_a_valid_config(torch_dev):
    import torch
    from fesom_jax_amd import autotune
    from fesom_jax_amd.batch import FctBatch
    insts = _instances(torch_dev, [(3, 48, 2, 31), (4, 70, 3, 33)])
    batch = FctBatch(insts)
    s = torch.cuda.current_stream(torch_dev)
    best, table = autotune.tune_batch(batch, s, reps=1, warmup=1, key="test-batch")
    n = 1
    for v in autotune.BATCH_SPACE.values():
        n *= len(v)
    assert len(table) == n and all(ms > 0 for _, ms in table)
    assert set(best) == set(autotune.BATCH_SPACE)
    again, table2 = autotune.tune_batch(batch, s, key="test-batch")
    assert again == best and table2 == []
    batch.close()


def test_autotuner_descent_on_a_plan(icosa4_mesh, icosa4_case, torch_dev):
    """Axis-by-axis descent over every axis (edge tiles, the adf_h level block, XCD
    ranges) on a single plan: the picked shape gives the default shape's bits."""
    import ctypes
    import torch
    from ctypes import c_void_p
    from fesom_jax_amd import _lib, autotune
    from fesom_jax_amd.tracer_adv import FctPlan
    m, c = icosa4_mesh, icosa4_case
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(torch_dev)
    plan = FctPlan(m, torch_dev, max_tracers=2)
    T = [t(x) for x in c["T"][:2]]
    To = [t(x) for x in c["T_old"][:2]]
    out = [torch.empty_like(x) for x in T]
    arr = lambda xs: (c_void_p * len(xs))(*[x.data_ptr() for x in xs])
    keep = (arr(T), arr(To), arr(out))
    P = lambda a: ctypes.cast(a, ctypes.POINTER(c_void_p))
    f = [t(c[k]) for k in ("uv", "w_e", "helem", "hnode", "hnode_new")]
    args = _lib.FctStepArgs(ntracers=2, dt=c["dt"], uv=f[0].data_ptr(), w_e=f[1].data_ptr(),
                            helem=f[2].data_ptr(), hnode=f[3].data_ptr(),
                            hnode_new=f[4].data_ptr(), T=P(keep[0]), T_old=P(keep[1]),
                            T_new=P(keep[2]), w_i=None, Z3d=None, zbar3=None)
    s = torch.cuda.current_stream(torch_dev)
    sp = c_void_p(s.cuda_stream)
    _lib.check(plan.lib.fct_forward(plan.ptr, ctypes.byref(args), sp), "fct_forward")
    ref = [x.clone() for x in out]
    best, table = autotune.tune_plan(plan, args, s, reps=1, warmup=1, search="descent")
    assert set(best) == set(autotune.SPACE) and len(table) >= 1 + sum(
        len(v) - 1 for v in autotune.SPACE.values())
    _lib.check(plan.lib.fct_forward(plan.ptr, ctypes.byref(args), sp), "fct_forward")
    torch.cuda.synchronize()
    assert all(torch.equal(x, y) for x, y in zip(out, ref))
    plan.close()
