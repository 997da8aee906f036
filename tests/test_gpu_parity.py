"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the same
seeded inputs. Contract: max relative error <= 1e-12 on every lane (FP64); the
kernels keep the oracle's operation and summation order, so the tests also record
how many lanes differ at all (expected: none)."""

import numpy as np
import pytest

from conftest import rel_err, sharp_case
from oracle import fct_numpy as O

pytestmark = pytest.mark.gpu

TOL = 1e-12



def _gpu_step(mesh, case, dev, *, j=None, path="auto", **kw):
    import torch
    from fesom_jax_amd.tracer_adv import advect_fct
    Ts = case["T"] if j is None else [case["T"][j]]
    Tos = case["T_old"] if j is None else [case["T_old"][j]]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    out = advect_fct(mesh, t(case["uv"]), t(case["w_e"]), t(case["helem"]), t(case["hnode"]),
                     t(case["hnode_new"]), [t(x) for x in Ts], [t(x) for x in Tos],
                     dt=case["dt"], path=path,
                     **{k: (None if v is None else t(v)) if k != "exch" else v
                        for k, v in kw.items()})
    torch.cuda.synchronize()
    return [o[0].cpu().numpy() for o in out]


def _oracle(mesh, case, j, **kw):
    return O.advect_one_fct(mesh, case["uv"], case["w_e"], case["helem"], case["hnode"],
                            case["hnode_new"], case["T"][j], case["T_old"][j], dt=case["dt"],
                            **kw)[0]


def _check(name, gpu, ref, mesh):
    err = rel_err(gpu, ref)
    ndiff = int(np.sum(gpu != ref))
    print(f"[parity] {name}: rel_err={err:.3e} differing_lanes={ndiff}/{gpu.size}", flush=True)
    assert np.all(np.isfinite(gpu[mesh.node_layer_mask]))
    assert err <= TOL, f"{name}: relative error {err:.3e} > {TOL}"
    return ndiff


def test_pi_two_tracers(pi_mesh, pi_case, torch_dev):
    outs = _gpu_step(pi_mesh, pi_case, torch_dev)
    for j, g in enumerate(outs):
        _check(f"pi tracer {j}", g, _oracle(pi_mesh, pi_case, j), pi_mesh)


def test_pi_limiter_active(pi_mesh, pi_case, torch_dev):
    case = sharp_case(pi_mesh, pi_case, 2.0)
    ref, _, parts = O.advect_one_fct(pi_mesh, case["uv"], case["w_e"], case["helem"],
                                     case["hnode"], case["hnode_new"], case["T"][0],
                                     case["T_old"][0], dt=case["dt"], return_parts=True)
    m = pi_mesh.node_layer_mask
    assert (parts["plus_fac"][m] < 1).mean() > 0.02, "limiter not active enough"
    g = _gpu_step(pi_mesh, case, torch_dev)[0]
    _check("pi limiter-active", g, ref, pi_mesh)


def test_pi_second_step_and_ale(pi_mesh, pi_case, torch_dev):
    rng = np.random.default_rng(5)
    m = pi_mesh.node_layer_mask
    T0 = pi_case["T"][0]
    T1 = O.advect_one_fct(pi_mesh, pi_case["uv"], pi_case["w_e"], pi_case["helem"],
                          pi_case["hnode"], pi_case["hnode_new"], T0, T0, dt=2700.0)[0]
    hnn = pi_case["hnode"] * np.where(m, 1.0 + 0.01 * rng.standard_normal(m.shape), 1.0)
    case = dict(pi_case, T=[T1], T_old=[T0], hnode_new=hnn)
    g = _gpu_step(pi_mesh, case, torch_dev)[0]
    _check("pi step 2 + hnode_new != hnode", g, _oracle(pi_mesh, case, 0), pi_mesh)


def test_pi_wsplit_and_zstar(pi_mesh, pi_case, torch_dev):
    rng = np.random.default_rng(6)
    m = pi_mesh.node_iface_mask
    w_i = np.where(m, 0.3 * pi_case["w_e"] * rng.random(m.shape), 0.0)
    g = _gpu_step(pi_mesh, pi_case, torch_dev, j=0, w_i=w_i)[0]
    _check("pi w_i", g, _oracle(pi_mesh, pi_case, 0, w_i=w_i), pi_mesh)
    zb3 = pi_mesh.zbar_3d_n * (1.0 + 0.001 * rng.random(pi_mesh.zbar_3d_n.shape))
    Z3d = np.concatenate([0.5 * (zb3[:, :-1] + zb3[:, 1:]), zb3[:, -1:]], axis=1)
    g = _gpu_step(pi_mesh, pi_case, torch_dev, j=0, Z3d=Z3d, zbar3=zb3)[0]
    _check("pi Z3d/zbar3", g, _oracle(pi_mesh, pi_case, 0, Z3d=Z3d, zbar3=zb3), pi_mesh)


def test_icosa_three_tracers(icosa4_mesh, icosa4_case, torch_dev):
    outs = _gpu_step(icosa4_mesh, icosa4_case, torch_dev)
    for j, g in enumerate(outs):
        _check(f"icosa-4 tracer {j}", g, _oracle(icosa4_mesh, icosa4_case, j), icosa4_mesh)


def test_exch_phased_equals_fused(pi_mesh, pi_case, torch_dev):
    import torch
    seen = []

    def exch_copy(f, kind):
        seen.append((kind, tuple(f.shape)))
        return f.clone()

    fused = _gpu_step(pi_mesh, pi_case, torch_dev, j=0)[0]
    phased = _gpu_step(pi_mesh, pi_case, torch_dev, j=0, exch=exch_copy)[0]
    N, E, L = pi_mesh.nod2D, pi_mesh.elem2D, pi_mesh.nl
    assert seen == [("nod", (N, L)), ("elem", (E, L, 2)), ("nod", (N, L)), ("nod", (N, L))]
    assert np.array_equal(fused, phased)


def test_exch_hook_sees_reference_fields(pi_mesh, pi_case, torch_dev):
    """The intermediates handed to exch equal the oracle's LO, tr_xy, fct_plus,
    fct_minus; an exch that modifies LO changes the result as in the oracle."""
    got = {}

    def grab(f, kind):
        got.setdefault(kind, []).append(f.detach().cpu().numpy().copy())
        return f

    _gpu_step(pi_mesh, pi_case, torch_dev, j=0, exch=grab)
    _, _, parts = O.advect_one_fct(pi_mesh, pi_case["uv"], pi_case["w_e"], pi_case["helem"],
                                   pi_case["hnode"], pi_case["hnode_new"], pi_case["T"][0],
                                   pi_case["T_old"][0], dt=2700.0, return_parts=True)
    assert rel_err(got["nod"][0], parts["LO"]) <= TOL
    assert rel_err(got["elem"][0], parts["tr_xy"]) <= TOL
    assert rel_err(got["nod"][1], parts["fct_plus"]) <= TOL
    assert rel_err(got["nod"][2], parts["fct_minus"]) <= TOL

    gstate = {"n": 0}

    def bump_once(f, kind):
        gstate["n"] += 1
        return f * 1.001 if gstate["n"] == 1 else f

    g = _gpu_step(pi_mesh, pi_case, torch_dev, j=0, exch=bump_once)[0]
    state = {"n": 0}

    def np_bump(f, kind):
        state["n"] += 1
        return f * 1.001 if state["n"] == 1 else f

    ref = _oracle(pi_mesh, pi_case, 0, exch=np_bump)
    _check("pi exch(LO)*1.001", g, ref, pi_mesh)


def test_run_twice_bitwise(icosa4_mesh, icosa4_case, torch_dev):
    a = _gpu_step(icosa4_mesh, icosa4_case, torch_dev)
    b = _gpu_step(icosa4_mesh, icosa4_case, torch_dev)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_dry_lanes_copy_input(pi_mesh, pi_case, torch_dev):
    m = pi_mesh.node_layer_mask
    T = np.where(m, pi_case["T"][0], 7.25)
    case = dict(pi_case, T=[T], T_old=[T])
    g = _gpu_step(pi_mesh, case, torch_dev)[0]
    assert np.array_equal(g[~m], T[~m])
    _check("pi nonzero dry lanes", g, _oracle(pi_mesh, case, 0), pi_mesh)


def test_numpy_in_numpy_out_and_T_old_identity(pi_mesh, pi_case, torch_dev):
    from fesom_jax_amd.tracer_adv import advect_one_fct
    c = pi_case
    T = c["T"][0]
    Tn, To = advect_one_fct(pi_mesh, c["uv"], c["w_e"], c["helem"], c["hnode"],
                            c["hnode_new"], T, c["T_old"][0], dt=c["dt"])
    assert isinstance(Tn, np.ndarray) and To is T
    _check("pi numpy boundary", Tn, _oracle(pi_mesh, c, 0), pi_mesh)


def test_bad_inputs_raise(pi_mesh, pi_case, torch_dev):
    import torch
    from fesom_jax_amd.tracer_adv import advect_one_fct
    c = pi_case
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(torch_dev)
    with pytest.raises(ValueError):
        advect_one_fct(pi_mesh, t(c["uv"]), t(c["w_e"]), t(c["helem"]), t(c["hnode"]),
                       t(c["hnode_new"]), t(c["T"][0]).float(), t(c["T_old"][0]))
    with pytest.raises(ValueError):
        advect_one_fct(pi_mesh, t(c["uv"][:, :-1]), t(c["w_e"]), t(c["helem"]), t(c["hnode"]),
                       t(c["hnode_new"]), t(c["T"][0]), t(c["T_old"][0]))


@pytest.mark.parametrize("which", ["pi", "icosa4", "sharp"])
def test_fused_equals_phased_bitwise(which, pi_mesh, pi_case, icosa4_mesh, icosa4_case,
                                     torch_dev):
    mesh, case = {"pi": (pi_mesh, pi_case), "icosa4": (icosa4_mesh, icosa4_case),
                  "sharp": (pi_mesh, sharp_case(pi_mesh, pi_case, 2.0))}[which]
    from fesom_jax_amd.tracer_adv import get_plan
    assert get_plan(mesh, torch_dev).fused_available, f"{which}: mesh fails the fused-path gate"
    fused = _gpu_step(mesh, case, torch_dev, path="fused")
    phased = _gpu_step(mesh, case, torch_dev, path="phased")
    for j, (f, p) in enumerate(zip(fused, phased)):
        ref = _oracle(mesh, case, j)
        _check(f"{which} fused tracer {j}", f, ref, mesh)
        assert np.array_equal(f, p), f"{which} tracer {j}: fused and phased differ"


def test_fused_zstar_and_many_tracers(icosa4_mesh, icosa4_case, torch_dev):
    """Z3d/zbar3 on the fused path, and 6 tracers (more than one workspace chunk)."""
    rng = np.random.default_rng(9)
    m = icosa4_mesh
    zb3 = m.zbar_3d_n * (1.0 + 0.001 * rng.random(m.zbar_3d_n.shape))
    Z3d = np.concatenate([0.5 * (zb3[:, :-1] + zb3[:, 1:]), zb3[:, -1:]], axis=1)
    g = _gpu_step(m, icosa4_case, torch_dev, j=0, path="fused", Z3d=Z3d, zbar3=zb3)[0]
    _check("icosa4 fused Z3d", g, _oracle(m, icosa4_case, 0, Z3d=Z3d, zbar3=zb3), m)
    Ts = [icosa4_case["T"][j % 3] * (1.0 + 0.01 * j) for j in range(6)]
    case = dict(icosa4_case, T=Ts, T_old=[x * 0.999 for x in Ts])
    outs = _gpu_step(m, case, torch_dev, path="fused")
    for j, g in enumerate(outs):
        _check(f"icosa4 fused 6 tracers [{j}]", g, _oracle(m, case, j), m)


def test_icosa_nl70_two_slots(torch_dev):
    """nl = 70 needs two 64-lane slots per column on the fused path."""
    from fesom_jax_amd.inputs import make_case
    from fesom_jax_amd.meshgen import icosahedral_mesh
    m = icosahedral_mesh(4, 70, seed=21)
    case = make_case(m, seed=22, dt=2700.0)
    assert int(m.nlevels_nod2D.max()) > 65 and int(m.nlevels_nod2D.min()) < 60
    outs = _gpu_step(m, case, torch_dev, path="fused")
    for j, g in enumerate(outs):
        _check(f"icosa4 nl70 fused [{j}]", g, _oracle(m, case, j), m)
    sh = sharp_case(m, case, 2.0)
    g = _gpu_step(m, sh, torch_dev, path="fused")[0]
    _check("icosa4 nl70 fused sharp", g, _oracle(m, sh, 0), m)


@pytest.mark.parametrize("shape", [(1, 0.85, 2), (1, 0.7, 1), (1, 0.7, 0), (2, 0.9, 2)],
                         ids=["one_triangle", "two_triangles", "six_triangles", "nine_triangles"])
def test_tiny_meshes(shape, torch_dev):
    """The smallest meshes: one triangle (three coastal edges, no up/down triangles), two
    triangles (one interior edge), columns of one to three layers (nl = 4), on the
    default and the phased path, against the oracle, smooth and with the limiter
    active."""
    from fesom_jax_amd.inputs import make_case
    from fesom_jax_amd.meshgen import icosahedral_mesh
    from fesom_jax_amd.tracer_adv import get_plan
    lev, lf, seed = shape
    m = icosahedral_mesh(lev, 4, seed=seed, land_fraction=lf, min_levels=2,
                         polar_cap_deg=60.0)
    c = make_case(m, seed=seed + 50, dt=2700.0)
    print(f"[parity] tiny mesh N={m.nod2D} E={m.elem2D} D={m.edge2D} "
          f"fused={get_plan(m, torch_dev).fused_available}", flush=True)
    for name, case in (("smooth", c), ("sharp", sharp_case(m, c, 2.0))):
        for path in ("auto", "phased"):
            outs = _gpu_step(m, case, torch_dev, path=path)
            for j, g in enumerate(outs):
                _check(f"tiny N={m.nod2D} {name} {path} [{j}]", g, _oracle(m, case, j), m)


def test_degree_eight_nl70_depth_split(torch_dev):
    """Nodes with 8 neighbours (8 table slots per node) at nl = 70: the node passes split
    by column depth (one-slot kernels for the columns that end above level 64) on 8-slot
    tables, against the oracle, smooth and with the limiter active, and against the
    phased path bit for bit."""
    from fesom_jax_amd.inputs import make_case
    from fesom_jax_amd.meshgen import icosahedral_mesh
    from fesom_jax_amd.tracer_adv import get_plan
    m = icosahedral_mesh(4, 70, seed=31, degree8=30)
    case = make_case(m, seed=32, dt=2700.0)
    assert np.bincount(m.edges.reshape(-1)).max() == 8
    assert int(m.nlevels_nod2D.max()) > 65 and int(m.nlevels_nod2D.min()) < 60
    assert get_plan(m, torch_dev).fused_available
    for name, c in (("smooth", case), ("sharp", sharp_case(m, case, 2.0))):
        outs = _gpu_step(m, c, torch_dev, path="fused")
        ph = _gpu_step(m, c, torch_dev, path="phased")
        for j, g in enumerate(outs):
            _check(f"icosa4 nl70 degree 8 {name} [{j}]", g, _oracle(m, c, j), m)
            assert np.array_equal(g, ph[j]), j


@pytest.mark.parametrize("knobs", [{"XCD": 8}, {"WAVES_PER_GROUP": 2, "XCD": 3},
                                   {"WAVES_PER_GROUP": 1, "VFLUX_EDGES": 2, "ADFH_EDGES": 1},
                                   {"XCD": 8, "XCD_CHUNK": 16}, {"XCD": 3, "XCD_CHUNK": 1000},
                                   {"PAIR_TRACERS": 0}, {"ADFH_CHUNK": 0},
                                   {"ADFH_CHUNK": 4, "ADFH_THREADS": 2},
                                   {"ADFH_CHUNK": 8, "ADFH_THREADS": 4}])
def test_launch_shapes_do_not_change_results(knobs, torch_dev):
    """Other launch shapes give the default launch's bits."""
    from fesom_jax_amd import _lib
    from fesom_jax_amd.inputs import make_case
    from fesom_jax_amd.meshgen import icosahedral_mesh
    from fesom_jax_amd.tracer_adv import get_plan
    m = icosahedral_mesh(4, 70, seed=23)
    case = sharp_case(m, make_case(m, seed=24, dt=2700.0), 2.0)
    ref = _gpu_step(m, case, torch_dev, path="fused")
    plan = get_plan(m, torch_dev)
    try:
        for k, v in knobs.items():
            _lib.check(plan.lib.fct_plan_set_tuning(plan.ptr, getattr(_lib, "FCT_TUNE_" + k), v),
                       "tuning")
        got = _gpu_step(m, case, torch_dev, path="fused")
    finally:
        for k, v in (("XCD", 8), ("XCD_CHUNK", 8), ("WAVES_PER_GROUP", 4), ("VFLUX_EDGES", 4),
                     ("ADFH_EDGES", 2), ("PAIR_TRACERS", 1), ("ADFH_CHUNK", 8),
                     ("ADFH_THREADS", 1)):
            plan.lib.fct_plan_set_tuning(plan.ptr, getattr(_lib, "FCT_TUNE_" + k), v)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("which", ["pi", "icosa4", "nl70", "nl70_deg8"])
def test_upwind_step_bitwise(which, pi_mesh, pi_case, icosa4_mesh, icosa4_case, torch_dev):
    """advect_one on the GPU against the oracle (tracer_adv.py:154-165), bit for bit."""
    import torch
    from fesom_jax_amd.inputs import make_case
    from fesom_jax_amd.meshgen import icosahedral_mesh
    from fesom_jax_amd.tracer_adv import advect_one
    if which == "pi":
        m, c = pi_mesh, pi_case
    elif which == "icosa4":
        m, c = icosa4_mesh, icosa4_case
    else:
        m = icosahedral_mesh(4, 70, seed=21, degree8=30 if which == "nl70_deg8" else 0)
        c = make_case(m, seed=22, dt=2700.0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(torch_dev)
    for j in range(len(c["T"])):
        To = 0.98 * c["T"][j] + 0.02 * c["T"][0]
        want, _ = O.advect_one(m, c["uv"], c["w_e"], c["helem"], c["hnode"], c["hnode_new"],
                               c["T"][j], To, dt=c["dt"])
        got, Tr = advect_one(m, t(c["uv"]), t(c["w_e"]), t(c["helem"]), t(c["hnode"]),
                             t(c["hnode_new"]), t(c["T"][j]), t(To), dt=c["dt"])
        torch.cuda.synchronize()
        assert np.array_equal(got.cpu().numpy(), want), f"{which} tracer {j}"


def test_small_icosa_matches_loop_oracle(torch_dev):
    """Whole path on a 162-node sphere: the GPU equals the numpy oracle bit for bit and
    the independent C loop restatement (oracle/fct_loops.c) to 1e-11, smooth and with
    the limiter active."""
    from fesom_jax_amd.inputs import make_case
    from fesom_jax_amd.meshgen import icosahedral_mesh
    from oracle import loops as LP
    m = icosahedral_mesh(2, 24, seed=21)
    c = make_case(m, seed=22, dt=2700.0)
    for case in (c, sharp_case(m, c, 3.0)):
        got = _gpu_step(m, case, torch_dev)
        for j in range(len(case["T"])):
            want = O.advect_one_fct(m, case["uv"], case["w_e"], case["helem"], case["hnode"],
                                    case["hnode_new"], case["T"][j], case["T_old"][j],
                                    dt=case["dt"])[0]
            loops = LP.advect_one_fct(m, case["uv"], case["w_e"], case["helem"],
                                      case["hnode"], case["hnode_new"], case["T"][j],
                                      case["T_old"][j], dt=case["dt"])
            assert np.array_equal(got[j], want), j
            assert rel_err(got[j], loops, m.node_layer_mask) <= 1e-11, j


@pytest.mark.parametrize("nl", [48, 70])
def test_shuffled_numbering(nl, torch_dev):
    """A mesh numbered at random (nodes, triangles, edges): the plan's own locality order
    drives the kernels, and the results are the oracle's bits on that numbering (every
    per-node sum still runs in the numbering's (edge, side) order)."""
    from fesom_jax_amd.inputs import make_case
    from fesom_jax_amd.meshgen import icosahedral_mesh, shuffled
    from fesom_jax_amd.tracer_adv import get_plan
    base = icosahedral_mesh(4, nl, seed=31, degree8=6)
    m = shuffled(base, seed=32)
    assert get_plan(m, torch_dev).patch_order and not get_plan(base, torch_dev).patch_order
    case = sharp_case(m, make_case(m, seed=33, dt=2700.0), 2.0)
    outs = _gpu_step(m, case, torch_dev, path="fused")
    for j, g in enumerate(outs):
        assert _check(f"shuffled nl={nl} tracer {j}", g, _oracle(m, case, j), m) == 0


def test_two_streams_share_a_plan(icosa4_mesh, icosa4_case, torch_dev):
    """Two torch streams stepping different states on one mesh share its cached plan
    (and workspace); the plan orders them, so each result keeps its single-stream bits."""
    import torch
    from fesom_jax_amd.tracer_adv import advect_fct
    m, c = icosa4_mesh, icosa4_case
    sharp = sharp_case(m, c, 2.0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(torch_dev)
    flow = [t(c[k]) for k in ("uv", "w_e", "helem", "hnode", "hnode_new")]
    flow2 = [t(sharp[k]) for k in ("uv", "w_e")] + flow[2:]
    A = ([t(x) for x in c["T"]], [t(x) for x in c["T_old"]])
    B = ([t(sharp["T"][0])], [t(sharp["T_old"][0])])
    refA = [o[0].clone() for o in advect_fct(m, *flow, *A, dt=c["dt"])]
    refB = [o[0].clone() for o in advect_fct(m, *flow2, *B, dt=c["dt"])]
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(torch_dev), torch.cuda.Stream(torch_dev)
    s1.wait_stream(torch.cuda.current_stream(torch_dev))
    s2.wait_stream(torch.cuda.current_stream(torch_dev))
    outs = []
    for _ in range(3):
        with torch.cuda.stream(s1):
            oa = advect_fct(m, *flow, *A, dt=c["dt"])
        with torch.cuda.stream(s2):
            ob = advect_fct(m, *flow2, *B, dt=c["dt"])
        outs.append((oa, ob))
    torch.cuda.synchronize()
    for oa, ob in outs:
        for x, y in zip(oa, refA):
            assert torch.equal(x[0], y)
        assert torch.equal(ob[0][0], refB[0])


def test_cloned_plan_shares_tables(icosa4_mesh, icosa4_case, torch_dev):
    """fct_plan_clone: the clone shares the mesh tables (same table id, its own bytes
    are workspace only), gives the original's bits, and keeps working after the
    original is destroyed (the tables go with the last plan)."""
    import torch
    from fesom_jax_amd.tracer_adv import FctPlan, advect_fct
    m, c = icosa4_mesh, icosa4_case
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(torch_dev)
    flow = [t(c[k]) for k in ("uv", "w_e", "helem", "hnode", "hnode_new")]
    Ts, Tos = [t(x) for x in c["T"]], [t(x) for x in c["T_old"]]
    base = FctPlan(m, torch_dev, max_tracers=4)
    clone = FctPlan(m, torch_dev, max_tracers=2, share=base)
    assert clone.tables() == base.tables() and clone.tables()[1] > 0
    assert clone.own_bytes() + clone.tables()[1] == clone.device_bytes()
    assert clone.own_bytes() < base.own_bytes()      # 2 tracer slots against 4
    ref = [o[0].clone() for o in advect_fct(m, *flow, Ts, Tos, dt=c["dt"], plan=base)]
    got = [o[0].clone() for o in advect_fct(m, *flow, Ts, Tos, dt=c["dt"], plan=clone)]
    torch.cuda.synchronize()
    base.close()
    again = [o[0] for o in advect_fct(m, *flow, Ts, Tos, dt=c["dt"], plan=clone)]
    torch.cuda.synchronize()
    for r, g, a in zip(ref, got, again):
        assert torch.equal(r, g) and torch.equal(r, a)
    with pytest.raises(ValueError):
        FctPlan(icosa4_mesh, torch_dev, share=FctPlan(_pi_like(), torch_dev))
    clone.close()


def _pi_like():
    from fesom_jax_amd.meshgen import icosahedral_mesh
    return icosahedral_mesh(2, 24, seed=3)


@pytest.mark.parametrize("nl, K", [(48, 2), (70, 3), (48, 4)])
def test_tracers_share_wavefronts(nl, K, torch_dev):
    """Shallow columns take the multi-tracer node passes (two tracer slots per
    wavefront, the last alone for odd K), for nl <= 64 and nl > 64; bit for bit with
    the oracle and with one tracer per wavefront (limiter-active state)."""
    from fesom_jax_amd import _lib
    from fesom_jax_amd.inputs import make_case
    from fesom_jax_amd.meshgen import icosahedral_mesh
    from fesom_jax_amd.tracer_adv import get_plan
    m = icosahedral_mesh(4, nl, seed=61)
    base = make_case(m, seed=62, dt=2700.0)
    sharp = sharp_case(m, base, 2.0)
    # the sharp tracer (limiter active) and smooth ones of the base case
    Ts = [sharp["T"][0]] + [base["T"][j % len(base["T"])] * (1.0 + 0.1 * j) for j in range(K - 1)]
    Tos = [sharp["T_old"][0]] + [base["T_old"][j % len(base["T"])] * (1.0 + 0.1 * j)
                                for j in range(K - 1)]
    case = dict(sharp, T=Ts, T_old=Tos)
    paired = _gpu_step(m, case, torch_dev, path="fused")
    for j, g in enumerate(paired):
        _check(f"nl{nl} K{K} shared tracer {j}", g, _oracle(m, case, j), m)
    plan = get_plan(m, torch_dev)
    try:
        _lib.check(plan.lib.fct_plan_set_tuning(plan.ptr, _lib.FCT_TUNE_PAIR_TRACERS, 0), "tuning")
        single = _gpu_step(m, case, torch_dev, path="fused")
    finally:
        plan.lib.fct_plan_set_tuning(plan.ptr, _lib.FCT_TUNE_PAIR_TRACERS, 1)
    for a, b in zip(paired, single):
        assert np.array_equal(a, b)
