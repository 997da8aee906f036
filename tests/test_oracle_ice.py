"""Pins for the sea-ice FCT oracle (oracle/ice_fct.py, restating
fesom_jax/ice_adv.py:37-194). The reference's own gate compares with a C dump of the
original kernel (fesom_jax/tests/test_ice_adv.py:49-56), which this checkout does not
hold; in its place: an independent element-loop restatement in C (oracle/fct_loops.c,
``loops_ice_fct``: scatter-adds over elements as the original kernel is structured),
total mass conserved (the low-order solve and the element antidiffusive fluxes both
conserve the area-weighted sum), the Zalesak bounds (fesom_jax/tests/
test_ice_adv.py:67-79 checks boundedness), a constant field at rest unchanged, and
planted mistakes that fail them."""

import copy
import math

import numpy as np
import pytest

from oracle import ice_fct as I
from oracle import loops as LP


def _case(mesh, seed=3, speed=0.5):
    from fesom_jax_amd.inputs import make_ice_case
    return make_ice_case(mesh, seed=seed, speed=speed)


def _solve(mod, cfg, mesh, c):
    f = I.fct_solve if mod == "numpy" else LP.ice_fct_solve
    return f(cfg, mesh, c["a_ice"], c["m_ice"], c["m_snow"], c["u_ice"], c["v_ice"])


def _with_shelf(mesh, frac=0.1, seed=5):
    """A copy of the mesh with ~frac of the elements under an ice shelf (ulevels 2) and
    the nodes all of whose elements are under it marked so too."""
    rng = np.random.default_rng(seed)
    m = copy.copy(mesh)
    ul = np.array(mesh.ulevels, copy=True)
    ul[rng.random(mesh.elem2D) < frac] = 2
    open_node = np.zeros(mesh.nod2D, bool)
    for q in range(3):
        open_node[mesh.elem_nodes[ul <= 1, q]] = True
    uln = np.where(open_node, 1, 2).astype(np.asarray(mesh.ulevels_nod2D).dtype)
    object.__setattr__(m, "ulevels", ul)
    object.__setattr__(m, "ulevels_nod2D", uln)
    return m


@pytest.mark.parametrize("dt", [500.0, 2.0e4])
def test_numpy_equals_c_loops(pi_mesh, dt):
    cfg = I.IceConfig(ice_dt=dt)
    c = _case(pi_mesh)
    for x, y in zip(_solve("numpy", cfg, pi_mesh, c), _solve("loops", cfg, pi_mesh, c)):
        assert np.array_equal(x, y)


def test_numpy_equals_c_loops_icosa_and_shelf(icosa4_mesh):
    cfg = I.IceConfig(ice_dt=2.0e4)
    c = _case(icosa4_mesh, seed=7)
    shelf = _with_shelf(icosa4_mesh)
    assert (shelf.ulevels_nod2D > 1).sum() > 0
    for m in (icosa4_mesh, shelf):
        for x, y in zip(_solve("numpy", cfg, m, c), _solve("loops", cfg, m, c)):
            assert np.array_equal(x, y)
    # nodes entirely under the shelf keep their values
    out = _solve("numpy", cfg, shelf, c)
    under = shelf.ulevels_nod2D > 1
    assert np.array_equal(out[0][under], c["a_ice"][under])


def test_mass_conserved_and_limiter_binds(pi_mesh):
    cfg = I.IceConfig(ice_dt=2.0e4)
    c = _case(pi_mesh)
    ar = np.asarray(pi_mesh.area)[:, 0]
    out = _solve("numpy", cfg, pi_mesh, c)
    for x, x0 in zip(out, (c["a_ice"], c["m_ice"], c["m_snow"])):
        m0, m1 = math.fsum(ar * x0), math.fsum(ar * x)
        assert abs(m1 - m0) <= 1e-14 * abs(m0)
        assert np.abs(x - x0).max() > 1e-3
    E = I.tg_entries(cfg, pi_mesh, c["u_ice"], c["v_ice"])
    rhs = I.tg_rhs(cfg, pi_mesh, E, c["a_ice"])
    _, parts = I.fem_fct(cfg, pi_mesh, c["a_ice"],
                         I.solve_low_order(cfg, pi_mesh, c["a_ice"], rhs),
                         I.solve_high_order(cfg, pi_mesh, rhs), return_parts=True)
    assert (parts["ae"] < 1).mean() > 0.1


def _bounds_violation(cfg, mesh, c, fem=None):
    E = I.tg_entries(cfg, mesh, c["u_ice"], c["v_ice"])
    worst = 0.0
    for X in (c["a_ice"], c["m_ice"], c["m_snow"]):
        rhs = I.tg_rhs(cfg, mesh, E, X)
        xl = I.solve_low_order(cfg, mesh, X, rhs)
        out, p = (fem or I.fem_fct)(cfg, mesh, X, xl, I.solve_high_order(cfg, mesh, rhs),
                                    return_parts=True)
        worst = max(worst, float((out - p["hi"]).max()), float((p["lo"] - out).max()))
    return worst


def test_zalesak_bounds(pi_mesh):
    cfg = I.IceConfig(ice_dt=2.0e4)
    assert _bounds_violation(cfg, pi_mesh, _case(pi_mesh)) <= 1e-14


def test_constant_at_rest_unchanged(pi_mesh):
    cfg = I.IceConfig(ice_dt=2.0e4)
    n = pi_mesh.nod2D
    c = dict(a_ice=np.full(n, 0.7), m_ice=np.full(n, 1.3), m_snow=np.full(n, 0.2),
             u_ice=np.zeros(n), v_ice=np.zeros(n))
    for x, v in zip(_solve("numpy", cfg, pi_mesh, c), (0.7, 1.3, 0.2)):
        assert np.abs(x - v).max() <= 1e-14


def test_planted_unlimited_flux_breaks_bounds(pi_mesh):
    """An oracle whose limiter passes every flux unscaled overshoots the bounds."""
    cfg = I.IceConfig(ice_dt=2.0e4)
    real = I.fem_fct

    def unlimited(cfg_, mesh, vals, vals_l, dvals, *, return_parts=False):
        out, p = real(cfg_, mesh, vals, vals_l, dvals, return_parts=True)
        from oracle.fct_numpy import scatter_add
        lim = np.where((np.asarray(mesh.ulevels) <= 1)[:, None], p["fluxes"], 0.0)
        out = vals_l + scatter_add(lim, np.asarray(mesh.elem_nodes), mesh.nod2D)
        return (out, p) if return_parts else out

    assert _bounds_violation(cfg, pi_mesh, _case(pi_mesh), fem=unlimited) > 1e-3


def test_planted_flux_sign_breaks_agreement(pi_mesh, monkeypatch):
    """Flipping the antidiffusive flux's sign in the oracle breaks the agreement with the
    C loops."""
    cfg = I.IceConfig(ice_dt=2.0e4)
    c = _case(pi_mesh)
    real = I.solve_high_order
    monkeypatch.setattr(I, "solve_high_order", lambda cfg_, mesh, rhs: -real(cfg_, mesh, rhs))
    x = _solve("numpy", cfg, pi_mesh, c)[0]
    y = _solve("loops", cfg, pi_mesh, c)[0]
    assert np.abs(x - y).max() > 1e-6
