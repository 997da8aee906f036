"""The three larger workloads at their configured sizes, through size-independent
properties (the oracle is too slow at these sizes):

* batched: the bench's draw of 64 instances (icosahedral levels 5-8, 47 or 70 levels,
  2-4 tracers): one batched step equals every instance's own plain step bit for bit,
  stays within each tracer's initial range, and conserves every instance's tracer mass
  when no flux crosses the surface;
* differentiation: 12 steps on icosahedral level 7 x 47 levels, 2 tracers, loss on the
  surface values of the first tracer: the gradient with respect to the second tracer is
  exactly zero, central finite differences along 20 seeded directions (kink-filtered:
  only directions whose two step sizes agree) match the hand-written adjoint, and a
  random output cotangent passes the dot-product identity;
* serving: a seeded draw of the serving stream (Poisson arrivals, icosahedral levels
  4-9, 47/70 levels, 2-4 tracers, 1-24 steps, 30 % adjoints) under a 200 GB budget:
  every result equals the request run alone bit for bit.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _free():
    import gc
    import torch
    gc.collect()
    torch.cuda.empty_cache()


def test_batched_64_instances(torch_dev):
    import torch
    from fesom_jax_amd.batch import FctBatch, Instance
    from fesom_jax_amd.inputs import make_case_device
    from fesom_jax_amd.meshgen import icosahedral_mesh
    _free()
    rng = np.random.default_rng(2024)
    specs = [(int(rng.integers(5, 9)), int(rng.choice([47, 70])), int(rng.integers(2, 5)))
             for _ in range(64)]
    print(f"[configs] batched: levels {sorted(set(s[0] for s in specs))}", flush=True)
    meshes, insts = {}, []
    for i, (lv, nl, K) in enumerate(specs):
        if (lv, nl) not in meshes:
            meshes[(lv, nl)] = icosahedral_mesh(lv, nl, seed=2024 + 10 * lv + nl)
        case = make_case_device(meshes[(lv, nl)], torch_dev, seed=3024 + i, dt=2700.0,
                                tracers=("T", "S", "T", "S")[:K])
        insts.append((meshes[(lv, nl)], Instance.from_case(meshes[(lv, nl)], case, torch_dev)))
        del case
    batch = FctBatch([inst for _, inst in insts])
    batch.prepare()
    T0 = [[t.clone() for t in inst.T()] for _, inst in insts]
    batch.step(graph=True)
    torch.cuda.synchronize()
    got = [[t.clone() for t in inst.T()] for _, inst in insts]
    for (m, inst), t0, g in zip(insts, T0, got):
        inst.phase = (inst.phase - 1) % 3
        inst.step_alone()
        torch.cuda.synchronize()
        wet = torch.from_numpy(m.node_layer_mask).to(torch_dev)
        for j, (a, b, x0) in enumerate(zip(g, inst.T(), t0)):
            assert torch.equal(a, b), ("batched != alone", j)
            lo, hi = float(x0[wet].min()), float(x0[wet].max())
            assert float(a[wet].min()) >= lo - 1e-11 and float(a[wet].max()) <= hi + 1e-11
    # no flux through the surface: every instance conserves each tracer's mass
    for (m, inst), t0 in zip(insts, T0):
        inst.phase = (inst.phase - 1) % 3
        top = torch.from_numpy((np.asarray(m.ulevels_nod2D) - 1).astype(np.int64)).to(torch_dev)
        inst.w_e[torch.arange(m.nod2D, device=torch_dev), top] = 0.0
    batch.prepare()
    batch.step(graph=False)
    torch.cuda.synchronize()
    worst = 0.0
    for (m, inst), t0 in zip(insts, T0):
        wet = torch.from_numpy(m.node_layer_mask).to(torch_dev)
        av = torch.from_numpy(np.asarray(m.areasvol)).to(torch_dev)
        for x0, x1 in zip(t0, inst.T()):
            m0 = torch.where(wet, av * inst.hnode * x0, 0.0).sum()
            m1 = torch.where(wet, av * inst.hnode_new * x1, 0.0).sum()
            worst = max(worst, float(abs(m1 - m0) / abs(m0)))
    print(f"[configs] batched: worst mass change {worst:.2e}", flush=True)
    assert worst <= 1e-13
    batch.close()
    del insts, batch
    _free()


@pytest.fixture(scope="module")
def adj_case(torch_dev):
    from fesom_jax_amd.inputs import make_case_device
    from fesom_jax_amd.meshgen import icosahedral_mesh
    _free()
    m = icosahedral_mesh(7, 47, seed=77)
    c = make_case_device(m, torch_dev, seed=78, dt=2700.0, tracers=("T", "S"))
    pert = make_case_device(m, torch_dev, seed=79, dt=2700.0, tracers=("S",))["T"][0]
    wet = __import__("torch").from_numpy(m.node_layer_mask).to(torch_dev)
    p = pert[wet]
    # target: initial T + 0.2 * a smoothed standard-normal field
    target = c["T"][0] + 0.2 * (pert - p.mean()) / p.std() * wet
    return m, c, target


def _loss_parts(m, c, T0, S0, uv, target, dev, nsteps=12):
    import torch
    from fesom_jax_amd.adjoint import rollout
    out = rollout(m, uv, c["w_e"], c["helem"], c["hnode"], c["hnode_new"], [T0, S0], [T0, S0],
                  nsteps, dt=c["dt"])
    surf = torch.from_numpy(m.node_layer_mask[:, 0]).to(dev)
    return out, torch.where(surf, out[0][:, 0] - target[:, 0], torch.zeros((), dtype=out[0].dtype,
                                                                             device=dev))


def test_adjoint_12_steps_icosa7(adj_case, torch_dev):
    import torch
    m, c, target = adj_case
    print(f"[configs] adjoint: icosa-7 N={m.nod2D} nl={m.nl}, 12 steps", flush=True)
    T0 = c["T"][0].clone().requires_grad_(True)
    S0 = c["T"][1].clone().requires_grad_(True)
    uv = c["uv"].clone().requires_grad_(True)
    _, r = _loss_parts(m, c, T0, S0, uv, target, torch_dev)
    loss = (r * r).sum()
    loss.backward()
    gT, gS, guv = T0.grad, S0.grad, uv.grad
    # salinity never meets temperature in the FCT
    assert int(torch.count_nonzero(gS)) == 0
    assert float(gT.abs().max()) > 0 and float(guv.abs().max()) > 0
    rng = np.random.default_rng(2)
    wet = m.node_layer_mask
    ewet = m.elem_layer_mask
    good, checked = 0, 0
    with torch.no_grad():
        def dloss(which, d, h):
            # loss(x + h d) - loss(x - h d), summed from per-lane differences
            args = lambda s: (T0 + s * d, S0, uv) if which == "T" else (T0, S0, uv + s * d)
            _, rp = _loss_parts(m, c, *args(h), target, torch_dev)
            _, rm = _loss_parts(m, c, *args(-h), target, torch_dev)
            return float(((rp - rm) * (rp + rm)).sum()) / (2 * h)
        for i in range(20):
            which = "T" if i % 2 == 0 else "uv"
            if which == "T":
                n = int(rng.integers(m.nod2D))
                d = np.zeros((m.nod2D, m.nl))
                d[n] = rng.standard_normal(m.nl) * wet[n]
                h, grad = 1e-4 * float(T0.abs().max()), gT
            else:
                e = int(rng.integers(m.elem2D))
                d = np.zeros((m.elem2D, m.nl, 2))
                d[e] = rng.standard_normal((m.nl, 2)) * ewet[e][:, None]
                h, grad = 1e-4 * float(uv.abs().max()), guv
            dt_ = torch.from_numpy(d).to(torch_dev)
            f1, f2 = dloss(which, dt_, h), dloss(which, dt_, 0.5 * h)
            ad = float((grad * dt_).sum())
            checked += 1
            if abs(f1 - f2) > 1e-5 * max(abs(f1), abs(f2), 1e-300):
                continue                     # a kink between the two step sizes
            fd = (4.0 * f2 - f1) / 3.0       # Richardson
            assert abs(fd - ad) <= 1e-6 * max(abs(fd), abs(ad)) + 1e-12, (i, which, fd, ad)
            good += 1
    print(f"[configs] adjoint: {good}/{checked} directions kink-free and matching", flush=True)
    assert good >= 12
    # dot-product identity <J v, w> = <v, J^T w> for a random output cotangent w and
    # directions v on one node column of each tracer (kink-filtered like the above)
    w = [torch.from_numpy(rng.standard_normal((m.nod2D, m.nl)) * wet).to(torch_dev)
         for _ in range(2)]
    T0.grad = S0.grad = uv.grad = None
    out, _ = _loss_parts(m, c, T0, S0, uv, target, torch_dev)
    (sum((o * x).sum() for o, x in zip(out, w))).backward()
    ok = 0
    with torch.no_grad():
        def jvw(vT, vS, h):
            op, _ = _loss_parts(m, c, T0 + h * vT, S0 + h * vS, uv, target, torch_dev)
            om, _ = _loss_parts(m, c, T0 - h * vT, S0 - h * vS, uv, target, torch_dev)
            return sum(float(((a - b) * x).sum()) for a, b, x in zip(op, om, w)) / (2 * h)
        for _ in range(6):
            vs = []
            for _k in range(2):
                n = int(rng.integers(m.nod2D))
                v = np.zeros((m.nod2D, m.nl))
                v[n] = rng.standard_normal(m.nl) * wet[n]
                vs.append(torch.from_numpy(v).to(torch_dev))
            h = 1e-4 * float(T0.abs().max())
            f1, f2 = jvw(*vs, h), jvw(*vs, 0.5 * h)
            if abs(f1 - f2) > 1e-5 * max(abs(f1), abs(f2), 1e-300):
                continue
            lhs = (4.0 * f2 - f1) / 3.0
            rhs = float((vs[0] * T0.grad).sum() + (vs[1] * S0.grad).sum())
            print(f"[configs] adjoint: <Jv, w> = {lhs:.12e}, <v, J^T w> = {rhs:.12e}", flush=True)
            assert abs(lhs - rhs) <= 1e-6 * max(abs(lhs), abs(rhs))
            ok += 1
    assert ok >= 2


def test_serving_stream_levels_4_to_9(torch_dev):
    import torch
    from fesom_jax_amd.serve import MeshCache, Server, fifo_loop, make_stream
    _free()
    specs = make_stream(200.0, 0.2, levels=(4, 9), nls=(47, 70), ks=(2, 3, 4), steps=(1, 24),
                        p_adjoint=0.3, seed=1)
    print(f"[configs] serving: {len(specs)} requests, levels "
          f"{sorted(set(s.level for s in specs))}, {sum(s.adjoint for s in specs)} adjoints",
          flush=True)
    assert any(s.level == 9 for s in specs) and any(s.adjoint for s in specs)
    meshes = MeshCache(seed=17)
    srv = Server(torch_dev, budget_bytes=200 << 30, max_active=16, meshes=meshes)
    summary = srv.run(specs)
    print(f"[configs] serving: peak held {summary['peak_bytes'] / 2**30:.1f} GiB, device peak "
          f"{torch.cuda.max_memory_reserved(torch_dev) / 2**30:.1f} GiB (torch); at the peak "
          f"{srv.peak_detail}", flush=True)
    srv.close()
    _free()
    assert summary["failed"] == 0 and summary["requests"] == len(specs)
    assert summary["peak_bytes"] <= 200 << 30
    alone = fifo_loop(specs, torch_dev, meshes=meshes)
    for s in specs:
        a, b = srv.results[s.rid], alone["results"][s.rid]
        assert set(a) == set(b), s.rid
        for key in a:
            xs = a[key] if isinstance(a[key], list) else [a[key]]
            ys = b[key] if isinstance(b[key], list) else [b[key]]
            for x, y in zip(xs, ys):
                assert np.array_equal(x, y), (s.rid, key)
    print(f"[configs] serving: p50 {summary['p50_s']:.2f} s, p99 {summary['p99_s']:.2f} s, "
          f"{summary['updates_per_s']:.3e} updates/s (FIFO {alone['updates_per_s']:.3e}); "
          f"wall {summary['wall_s']:.1f} s (FIFO {alone['wall_s']:.1f} s), host time "
          f"{summary['timing_s']}", flush=True)
    _free()
