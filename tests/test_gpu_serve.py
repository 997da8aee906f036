"""Serving: every request's result is the same served in the mix (continuous batching)
as run alone through the plain per-request loop, bit for bit."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone(torch_dev):
    from fesom_jax_amd.serve import MeshCache, Server, fifo_loop, make_stream
    specs = make_stream(200.0, 0.08, levels=(3, 4), nls=(47, 70), steps=(1, 6),
                        p_adjoint=0.4, seed=3)
    assert len(specs) >= 8 and any(s.adjoint for s in specs)
    meshes = MeshCache(seed=11)
    srv = Server(torch_dev, max_active=5, meshes=meshes)
    summary = srv.run(specs)
    assert summary["requests"] == len(specs)
    alone = fifo_loop(specs, torch_dev, meshes=meshes)
    for s in specs:
        a, b = srv.results[s.rid], alone["results"][s.rid]
        assert set(a) == set(b)
        for key in a:
            xs = a[key] if isinstance(a[key], list) else [a[key]]
            ys = b[key] if isinstance(b[key], list) else [b[key]]
            for x, y in zip(xs, ys):
                assert np.array_equal(x, y), (s.rid, key)
    assert summary["p99_s"] >= summary["p50_s"] > 0


def test_admission_respects_budget(torch_dev):
    from fesom_jax_amd.serve import MeshCache, Server, estimate_bytes, make_stream
    meshes = MeshCache(seed=11)
    specs = make_stream(500.0, 0.03, levels=(3, 3), nls=(47,), steps=(2, 3), p_adjoint=0.0,
                        seed=5)
    m = meshes.get(3, 47)
    # no two requests fit together; a request larger than the budget still runs alone
    one = min(estimate_bytes(m, s.K, s.nsteps, s.adjoint) for s in specs)
    srv = Server(torch_dev, budget_bytes=int(1.5 * one), max_active=8, meshes=meshes)
    peak = {"n": 0}
    orig = srv._start

    def start(spec, now):
        a = orig(spec, now)
        peak["n"] = max(peak["n"], len(srv.active) + 1)
        return a
    srv._start = start
    srv.run(specs)
    assert len(srv.latency) == len(specs)
    assert peak["n"] <= 1


def test_failures_stay_with_their_request(torch_dev):
    """A request with a NaN in its tracer and one whose setup the library refuses fail
    alone; every other request of the mix equals its stand-alone run bit for bit."""
    import dataclasses
    from fesom_jax_amd.serve import MeshCache, Server, fifo_loop, make_stream
    specs = make_stream(300.0, 0.05, levels=(3, 4), nls=(47, 70), steps=(2, 5),
                        p_adjoint=0.3, seed=8)
    assert len(specs) >= 8
    specs[1] = dataclasses.replace(specs[1], inject="nan")
    specs[4] = dataclasses.replace(specs[4], inject="fail")
    meshes = MeshCache(seed=12)
    srv = Server(torch_dev, max_active=6, meshes=meshes)
    # every result comes back through the staged download (large results do), in
    # chunks smaller than the fields
    srv.pinned_bytes = 0
    srv.stage_elems = 1 << 16
    summary = srv.run(specs)
    assert set(srv.errors) == {specs[1].rid, specs[4].rid}, srv.errors
    assert "non-finite" in srv.errors[specs[1].rid]
    assert "setup failed" in srv.errors[specs[4].rid]
    assert summary["failed"] == 2 and summary["requests"] == len(specs) - 2
    clean = [s for s in specs if s.inject is None]
    alone = fifo_loop(clean, torch_dev, meshes=meshes)
    for s in clean:
        a, b = srv.results[s.rid], alone["results"][s.rid]
        for key in b:
            xs = a[key] if isinstance(a[key], list) else [a[key]]
            ys = b[key] if isinstance(b[key], list) else [b[key]]
            for x, y in zip(xs, ys):
                assert np.array_equal(x, y), (s.rid, key)
    srv.close()


def test_budget_counts_every_live_plan(torch_dev):
    """The held bytes are the pooled plans' own counts (fct_plan_get_info), adjoint
    workspace included, with each mesh's shared tables counted once, and never pass
    the budget while every request fits it."""
    from fesom_jax_amd.serve import MeshCache, Server, estimate_bytes, make_stream
    meshes = MeshCache(seed=13)
    specs = make_stream(200.0, 0.06, levels=(3, 4), nls=(47,), steps=(2, 4), p_adjoint=0.5,
                        seed=9)
    big = max(estimate_bytes(meshes.get(s.level, s.nl), s.K, s.nsteps, s.adjoint) for s in specs)
    budget = int(2.5 * big)
    srv = Server(torch_dev, budget_bytes=budget, max_active=8, meshes=meshes)
    summary = srv.run(specs)
    assert not srv.errors and len(srv.latency) == len(specs)
    assert "over_budget" not in summary, summary.get("over_budget")
    pooled = [p for ps in srv.pool.values() for p in ps] + list(srv.adj_plans.values())
    assert pooled
    tables = dict(p.tables() for p in pooled)      # each mesh's tables once
    assert srv.held_bytes([]) == (sum(p.own_bytes() for p in pooled) + sum(tables.values())
                                  + sum(srv.base_bytes.values()))
    # plans of one mesh share one table set, the mesh's adjoint plan included
    per_mesh = {}
    for key, ps in srv.pool.items():
        for p in ps:
            per_mesh.setdefault(key[:2], set()).add(p.tables()[0])
    for key, p in srv.adj_plans.items():
        per_mesh.setdefault(key, set()).add(p.tables()[0])
    assert all(len(ids) == 1 for ids in per_mesh.values()), per_mesh
    assert any(s.adjoint for s in specs)
    assert srv.peak_bytes <= budget
    srv.close()
