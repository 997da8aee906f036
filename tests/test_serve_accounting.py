"""Host-side accounting of the server's HBM budget (no GPU): what an admission may
release to make room."""

import collections

from fesom_jax_amd import serve


class _FakePlan:
    def __init__(self, nbytes, table=None):
        self.nbytes = nbytes
        self.table = table or (id(self), 0)     # (table id, table bytes)
        self.closed = False

    def own_bytes(self):
        return self.nbytes

    def tables(self):
        return self.table

    def close(self):
        self.closed = True


class _Spec:
    def __init__(self, level, nl, adjoint=False):
        self.level, self.nl, self.adjoint = level, nl, adjoint


class _Active:
    def __init__(self, level, nl, nbytes):
        self.spec = _Spec(level, nl)
        self.bytes = nbytes


def _server(budget):
    srv = serve.Server("cpu", budget_bytes=budget)
    return srv


def test_release_keeps_the_plan_the_admission_reuses(monkeypatch):
    """An admission that will reuse an idle plan counts no plan bytes in its need, so
    that plan must survive the release: closing it would make the admission build a new
    plan the budget never saw."""
    srv = _server(1000)
    reused, other = _FakePlan(400), _FakePlan(300)
    srv.pool[(4, 47, 2)].append(reused)
    srv.pool[(5, 47, 2)].append(other)
    srv._refresh(reused)
    srv._refresh(other)
    ok = srv._release_idle(500, [], keep=reused)
    assert other.closed and not reused.closed
    assert ok            # 400 held + 500 needed fit 1000
    srv2 = _server(800)
    a, b = _FakePlan(400), _FakePlan(300)
    srv2.pool[(4, 47, 2)].append(a)
    srv2.pool[(5, 47, 2)].append(b)
    srv2._refresh(a)
    srv2._refresh(b)
    assert not srv2._release_idle(500, [], keep=a)    # only `a` could make room
    assert not a.closed


def test_release_keeps_the_admissions_base_inputs(monkeypatch):
    dropped = []
    monkeypatch.setattr(serve, "_drop_base_case", lambda mesh, dev, dt: dropped.append(mesh))
    srv = _server(1000)
    srv.meshes = type("M", (), {"get": staticmethod(lambda level, nl: (level, nl))})()
    srv.base_bytes = {(9, 70): 600, (8, 70): 300}
    ok = srv._release_idle(500, [_Active(8, 70, 0)], keep_base=(9, 70))
    # (9, 70) is the asking admission's, (8, 70) an active request's: neither goes
    assert dropped == [] and not ok
    ok = srv._release_idle(500, [], keep_base=(8, 70))
    assert dropped == [(9, 70)] and ok


def test_shared_tables_count_once_and_go_with_their_last_plan():
    srv = _server(1 << 40)
    a, b = _FakePlan(10, (77, 1000)), _FakePlan(20, (77, 1000))
    srv.pool[(6, 70, 2)].append(a)
    srv.pool[(6, 70, 3)].append(b)
    srv._refresh(a)
    srv._refresh(b)
    assert srv.held_bytes([]) == 10 + 20 + 1000
    srv.budget = 1000 + 20 + 5
    assert srv._release_idle(5, [], keep=b)        # closing `a` frees its 10 bytes only
    assert a.closed and srv.held_bytes([]) == 20 + 1000
    srv.budget = 10
    srv._release_idle(5, [])
    assert b.closed and srv.held_bytes([]) == 0 and not srv.table_sets


def test_held_bytes_sums_every_part():
    srv = _server(1 << 40)
    p = _FakePlan(7)
    srv._refresh(p)
    srv.base_bytes = {(3, 47): 11}
    srv._finishing = [(_Active(3, 47, 13), None, None, None)]
    assert srv.held_bytes([_Active(3, 47, 17)]) == 7 + 11 + 13 + 17
    assert isinstance(srv.timing, collections.Counter)


def test_adjoint_plan_stays_while_its_mesh_has_an_adjoint_in_flight():
    """The per-mesh adjoint plan is idle (releasable) only when no active or finishing
    request of that mesh asks for the adjoint."""
    srv = _server(1000)
    adj = _FakePlan(600)
    srv.adj_plans[(5, 47)] = adj
    srv._refresh(adj)
    a = _Active(5, 47, 100)
    a.spec.adjoint = True
    assert not srv._release_idle(500, [a])
    assert not adj.closed
    assert srv._release_idle(500, [_Active(5, 47, 100)])
    assert adj.closed and (5, 47) not in srv.adj_plans


def test_forward_admission_may_release_its_meshs_adjoint_plan():
    """An idle adjoint plan is kept for an admission of its mesh only when that
    admission runs an adjoint; a forward-only request of the mesh may release it (kept,
    it would push the held bytes past the budget)."""
    srv = _server(1000)
    adj = _FakePlan(600)
    srv.adj_plans[(9, 70)] = adj
    srv._refresh(adj)
    assert not srv._release_idle(500, [], keep_adj=(9, 70))
    assert not adj.closed
    assert srv._release_idle(500, [], keep_base=(9, 70))
    assert adj.closed and (9, 70) not in srv.adj_plans
    assert srv.held_bytes([]) + 500 <= srv.budget


def test_need_counts_the_tables_a_release_takes_away():
    """Closing the last live plan of the asking admission's mesh drops that mesh's
    tables from the held bytes, and the admission will build them again: the need is
    evaluated after each release, so the admission fits only if the tables fit too."""
    srv = _server(650)
    other_k = _FakePlan(300, (1, 200))          # the mesh's only live plan (K = 3)
    srv.pool[(9, 70, 3)].append(other_k)
    srv._refresh(other_k)

    def need():                                 # fields + plan for K = 2 (+ tables)
        return 500 + (200 if srv._live_plan(9, 70) is None else 0)
    assert not srv._release_idle(need, [])
    assert other_k.closed
    srv2 = _server(700)
    p = _FakePlan(300, (1, 200))
    srv2.pool[(9, 70, 3)].append(p)
    srv2._refresh(p)
    assert srv2._release_idle(lambda: 500 + (200 if srv2._live_plan(9, 70) is None else 0), [])
    assert srv2.held_bytes([]) == 0


def test_adjoint_plan_shares_the_admitted_requests_tables(monkeypatch):
    """The adjoint plan built for an admission shares the tables of that request's own
    plan (which is not yet among the active requests); a plan of its own would hold a
    second copy of the mesh's tables that the admission never counted."""
    from fesom_jax_amd import tracer_adv
    made = []

    class Plan(_FakePlan):
        def __init__(self, mesh, device, max_tracers=4, share=None):
            super().__init__(1, share.tables() if share is not None else (object(), 50))
            made.append(share)
    monkeypatch.setattr(tracer_adv, "FctPlan", Plan)
    mesh = type("Mesh", (), {"nod2D": 10, "elem2D": 16, "edge2D": 25, "nl": 5})()
    srv = _server(1 << 40)
    own = _FakePlan(7, (123, 50))
    srv._refresh(own)
    adj = srv._adjoint_plan(mesh, _Spec(9, 70, adjoint=True), share=own)
    assert made == [own] and adj.tables() == own.tables()
    assert len(srv.table_sets) == 1 and srv.held_bytes([]) == 7 + 50 + srv.plan_bytes[id(adj)]


def test_over_budget_snapshot_is_plain_data():
    """The snapshot of the first over-budget moment names every held part and the last
    admissions, and goes into the JSON summary as it is."""
    import json
    srv = _server(100)
    idle, adj = _FakePlan(30, (1, 20)), _FakePlan(40, (1, 20))
    srv.pool[(5, 47, 2)].append(idle)
    srv.adj_plans[(5, 47)] = adj
    srv._refresh(idle)
    srv._refresh(adj)
    srv.base_bytes = {(5, 47): 15}
    a = _Active(5, 47, 10)
    a.spec.rid = 7
    own = _FakePlan(25, (1, 20))
    srv._refresh(own)
    a.inst = type("Inst", (), {"plan": own})()
    srv._admissions.append((7, 5, 47, 2, False, 10, 105, 115))
    snap = srv._snapshot([a], srv.held_bytes([a]))
    assert snap["held"] == 30 + 40 + 25 + 20 + 15 + 10
    assert snap["plans"] == {"idle (5, 47, 2)": 30, "adjoint (5, 47)": 40, "active 7": 25}
    assert snap["active"] == {7: 10} and snap["admissions"][-1][0] == 7
    json.loads(json.dumps(snap))


def test_base_inputs_go_with_their_mesh(monkeypatch):
    """The base-input cache never hands a mesh another mesh's inputs, even when the
    other mesh's id is reused, and drops an entry when its mesh is collected."""
    import gc
    from fesom_jax_amd import inputs

    class M:
        pass
    built = []
    monkeypatch.setattr(inputs, "make_case_device",
                        lambda mesh, device, **kw: built.append(id(mesh)) or {"mesh": id(mesh)})
    a = M()
    assert serve._base_case(a, "cpu", 1.0) == {"mesh": id(a)}
    assert serve._has_base_case(a, "cpu", 1.0)
    key = (id(a), "cpu", 1.0)
    del a
    gc.collect()
    assert key not in serve._BASE
    b = M()
    assert not serve._has_base_case(b, "cpu", 1.0) or serve._BASE[(id(b), "cpu", 1.0)][0]() is b
    serve._base_case(b, "cpu", 1.0)
    assert serve._base_case(b, "cpu", 1.0) == {"mesh": id(b)} and len(built) == 2


def test_staged_copy_covers_every_chunk():
    """The staged download copies every element, whatever the size relation between the
    field and the staging buffer (CPU tensors stand for device and pinned memory)."""
    import torch
    waits = []
    for n, cs in ((1, 7), (7, 7), (20, 7), (21, 7), (5, 64)):
        src = torch.arange(n, dtype=torch.float64) * 0.5 - 3.0
        stage = torch.full((cs,), float("nan"), dtype=torch.float64)
        out = serve.staged_to_host(src, stage, lambda: waits.append(1))
        assert torch.equal(out, src) and out.data_ptr() != src.data_ptr()
    assert len(waits) == 1 + 1 + 3 + 3 + 1
