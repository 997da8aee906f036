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
