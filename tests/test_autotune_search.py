"""The autotuner's search logic on the host (no GPU): timing replaced by a cost of the
applied knobs."""

from fesom_jax_amd import autotune


def _fake(monkeypatch, cost):
    state = {}
    monkeypatch.setattr(autotune, "time_ms", lambda run, stream, reps=3, warmup=1: cost(state))
    return state, (lambda k, v: state.__setitem__(k, v))


def test_descent_finds_the_separable_optimum(monkeypatch):
    space = {1: (1, 8), 5: (0, 8, 32), 2: (1, 2, 4)}
    want = {1: 8, 5: 32, 2: 2}
    state, set_knob = _fake(monkeypatch, lambda st: sum(abs(st[k] - want[k]) for k in want))
    best, table = autotune.tune(set_knob, None, None, space=space, search="descent")
    assert best == want and state == want
    # a sweep that finds it, then a sweep that changes nothing
    assert len(table) <= 1 + 2 * sum(len(v) for v in space.values())


def test_grid_times_every_configuration(monkeypatch):
    space = {3: (2, 4), 4: (1, 2)}
    state, set_knob = _fake(monkeypatch, lambda st: st[3] * 10 + st[4])
    best, table = autotune.tune(set_knob, None, None, space=space)
    assert len(table) == 4 and best == {3: 2, 4: 1} and state == best


def test_descent_starts_from_the_library_defaults(monkeypatch):
    state, set_knob = _fake(monkeypatch, lambda st: 1.0)      # flat: nothing beats it
    best, table = autotune.tune(set_knob, None, None, space=autotune.SPACE, search="descent")
    assert best == {k: autotune.DEFAULT[k] for k in autotune.SPACE}
    assert autotune.BATCH_SPACE.keys() <= autotune.SPACE.keys()
