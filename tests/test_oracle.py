"""Pins for the CPU oracle (oracle/fct_numpy.py).

None of these retypes the oracle's formulas: each check compares against brute force,
a textbook routine, a closed form or an invariant, and each "planted" test shows the
check fails when a plausible mistake is put into the oracle (monkeypatched on the CPU
for that test only).
"""

import math

import numpy as np
import pytest

from conftest import rel_err, sharp_case
from oracle import fct_numpy as O


# --------------------------------------------------------------------------
# primitives (fesom_jax/tests/test_ops.py:45-106, 161-176 of the reference)
# --------------------------------------------------------------------------
def test_scatter_add_matches_loop_and_masks_sentinel():
    rng = np.random.default_rng(0)
    seg = rng.integers(-1, 7, size=(40, 2))
    vals = rng.standard_normal((40, 2, 3))
    got = O.scatter_add(vals, seg, 7)
    ref = np.zeros((7, 3))
    for i in range(40):
        for s in range(2):
            if seg[i, s] >= 0:
                ref[seg[i, s]] += vals[i, s]
    assert np.array_equal(got, ref)           # same sequential order -> same bits
    with pytest.raises(ValueError):
        O.scatter_add(vals[:10], seg, 7)


def test_scatter_transpose_is_gather():
    """<scatter(v), w> == <v, gather(w)> (the VJP identity of fesom_jax/ops.py:12-17)."""
    rng = np.random.default_rng(1)
    seg = rng.integers(0, 9, size=(50,))
    v = rng.standard_normal(50)
    w = rng.standard_normal(9)
    assert math.isclose(float(O.scatter_add(v, seg, 9) @ w), float(v @ w[seg]), rel_tol=1e-13)


def test_segment_extrema_empty_segments():
    vals = np.array([[1.0], [5.0], [-2.0]])
    seg = np.array([0, 0, 2])
    assert O.segment_max(vals, seg, 4)[:, 0].tolist() == [5.0, -np.inf, -2.0, -np.inf]
    assert O.segment_min(vals, seg, 4)[:, 0].tolist() == [1.0, np.inf, -2.0, np.inf]


def test_tdma_matches_linalg_solve():
    rng = np.random.default_rng(2)
    for K in (1, 2, 5, 17):
        a = rng.standard_normal((4, K)); c = rng.standard_normal((4, K))
        b = 4.0 + np.abs(a) + np.abs(c)
        d = rng.standard_normal((4, K))
        x = O.tdma(a, b, c, d)
        for r in range(4):
            M = np.diag(b[r]) + np.diag(a[r, 1:], -1) + np.diag(c[r, :-1], 1)
            assert np.allclose(M @ x[r], d[r], rtol=1e-12, atol=1e-12)


# --------------------------------------------------------------------------
# gradient stencils: exact for linear fields (a7-a9)
# --------------------------------------------------------------------------
def test_element_gradient_exact_for_linear_field(pi_mesh):
    """gradient_sca contracts a field linear in (x*cos, y) metres to its gradient."""
    m = pi_mesh
    lon, lat = m.coord_nod2D[:, 0], m.coord_nod2D[:, 1]
    gx, gy = 3.0e-6, -2.0e-6
    T = np.zeros((m.nod2D, m.nl))
    en = m.elem_nodes
    # per element, a field linear in the element's local metric coordinates
    el = 100
    x = (lon - lon[en[el, 0]]) * m.elem_cos[el] * 6367500.0
    y = (lat - lat[en[el, 0]]) * 6367500.0
    T[:, :] = (gx * x + gy * y)[:, None]
    tr = O.tracer_gradient_elements(m, T)
    assert np.allclose(tr[el, 0], [gx, gy], rtol=1e-6, atol=1e-12)


def test_node_average_of_constant_gradient(pi_mesh):
    m = pi_mesh
    tr = np.zeros((m.elem2D, m.nl, 2))
    tr[..., 0] = np.where(m.elem_layer_mask, 2.5, 0.0)
    tr[..., 1] = np.where(m.elem_layer_mask, -1.5, 0.0)
    g = O.node_avg_grad(m, tr)
    wet = m.node_layer_mask
    assert np.allclose(g[..., 0][wet], 2.5, rtol=1e-14)
    assert np.allclose(g[..., 1][wet], -1.5, rtol=1e-14)


# --------------------------------------------------------------------------
# whole-step invariants
# --------------------------------------------------------------------------
def _step(m, c, T, Told, **kw):
    return O.advect_one_fct(m, c["uv"], c["w_e"], c["helem"], c["hnode"], c["hnode_new"],
                            T, Told, dt=c["dt"], **kw)


def test_constant_tracer_preserved(pi_mesh, pi_case):
    """fesom_jax/tests/test_tracers.py:599-606: a constant tracer stays constant."""
    m = pi_mesh
    Tc = np.where(m.node_layer_mask, 20.0, 0.0)
    out, told = _step(m, pi_case, Tc, Tc)
    assert told is Tc
    assert np.max(np.abs(out[m.node_layer_mask] - 20.0)) < 1e-11


def test_constant_tracer_preserved_icosa(icosa4_mesh, icosa4_case):
    m = icosa4_mesh
    Tc = np.where(m.node_layer_mask, 35.0, 0.0)
    out, _ = _step(m, icosa4_case, Tc, Tc)
    assert np.max(np.abs(out[m.node_layer_mask] - 35.0)) < 1e-11


def _mass(m, h, T):
    return math.fsum((m.areasvol * h * T)[m.node_layer_mask].tolist())


def test_mass_budget_closes(pi_mesh, pi_case):
    """Change of sum(areasvol * h * T) equals dt * the surface boundary flux (LO
    upwind + limited antidiffusive surface flux), to 1e-14 relative."""
    m, c = pi_mesh, pi_case
    for T, Told in ((c["T"][0], c["T"][0]), (c["T"][1], 0.98 * c["T"][1])):
        out, _, parts = _step(m, c, T, Told, return_parts=True)
        nz = (m.ulevels_nod2D - 1).astype(int)
        idx = np.arange(m.nod2D)
        surf = parts["flux_v_lo"][idx, nz] + parts["adf_v_lim"][idx, nz]
        budget = c["dt"] * math.fsum(surf.tolist())
        lhs = _mass(m, c["hnode_new"], out) - _mass(m, c["hnode"], T)
        assert abs(lhs - budget) <= 1e-14 * _mass(m, c["hnode"], T)


def test_zalesak_bounds_and_factor_range(pi_mesh, pi_case):
    """With the limiter active, LO + fct_ttf_min <= T_new <= LO + fct_ttf_max and
    R+/R- lie in [0, 1] (to rounding)."""
    m = pi_mesh
    c = sharp_case(m, pi_case, 2.0)
    out, _, p = _step(m, c, c["T"][0], c["T_old"][0], return_parts=True)
    w = m.node_layer_mask
    assert (p["plus_fac"][w] < 1.0).mean() > 0.02
    up = p["LO"] + p["fct_ttf_max"]
    dn = p["LO"] + p["fct_ttf_min"]
    assert np.all(out[w] <= up[w] + 1e-12) and np.all(out[w] >= dn[w] - 1e-12)
    for f in (p["plus_fac"], p["minus_fac"]):
        assert np.all(f[w] <= 1.0) and np.all(f[w] >= -1e-15)


def test_lo_and_gradients_use_values_ho_uses_ab2(pi_mesh, pi_case):
    """LO and the element gradients are built from T, the high-order fluxes from the
    AB2 field (fesom_jax/tracer_adv.py:549-568): changing T_old leaves LO and tr_xy
    bit-identical and changes adf_h and adf_v."""
    m, c = pi_mesh, pi_case
    T = c["T"][0]
    _, _, p1 = _step(m, c, T, T, return_parts=True)
    _, _, p2 = _step(m, c, T, 0.9 * T, return_parts=True)
    assert np.array_equal(p1["LO"], p2["LO"])
    assert np.array_equal(p1["tr_xy"], p2["tr_xy"])
    assert not np.array_equal(p1["adf_h"], p2["adf_h"])
    assert not np.array_equal(p1["adf_v"], p2["adf_v"])


def test_lo_and_antidiffusive_fluxes_linear_in_tracer(pi_mesh, pi_case):
    """LO, adf_h and adf_v are linear in (T, T_old) for a fixed flow: second
    differences along a direction vanish to rounding."""
    m, c = pi_mesh, pi_case
    T = c["T"][0]
    dT = 1e-3 * np.where(m.node_layer_mask, np.sin(np.arange(m.nod2D))[:, None], 0.0)
    parts = [_step(m, c, T + s * dT, T + s * dT, return_parts=True)[2] for s in (0, 1, 2)]
    for key in ("LO", "adf_h", "adf_v"):
        a, b, d = (p[key] for p in parts)
        scale = np.max(np.abs(b)) + 1e-300
        assert np.max(np.abs(d - 2 * b + a)) <= 1e-12 * scale, key


# --------------------------------------------------------------------------
# planted mistakes: each of these must make one of the checks above fail
# --------------------------------------------------------------------------
def test_planted_sign_error_in_vflux_breaks_constant_preservation(pi_mesh, pi_case,
                                                                  monkeypatch):
    orig = O.edge_vflux
    monkeypatch.setattr(O, "edge_vflux", lambda m, uv, h: -orig(m, uv, h))
    m = pi_mesh
    Tc = np.where(m.node_layer_mask, 20.0, 0.0)
    out, _ = _step(m, pi_case, Tc, Tc)
    assert np.max(np.abs(out[m.node_layer_mask] - 20.0)) > 1e-6


def test_planted_dropped_vertical_term_breaks_mass_budget(pi_mesh, pi_case, monkeypatch):
    orig = O._shift_up
    m, c = pi_mesh, pi_case

    def bad_shift_up(x):            # loses the bottom-face flux of every layer
        return np.zeros_like(x) if x.shape == (m.nod2D, m.nl) else orig(x)
    monkeypatch.setattr(O, "_shift_up", bad_shift_up)
    Tc = np.where(m.node_layer_mask, 20.0, 0.0)
    out, _ = _step(m, c, Tc, Tc)
    assert np.max(np.abs(out[m.node_layer_mask] - 20.0)) > 1e-9


def test_planted_transposed_limiter_breaks_bounds(pi_mesh, pi_case, monkeypatch):
    m = pi_mesh
    c = sharp_case(m, pi_case, 2.0)
    orig = O.zalesak_limit

    def swapped(mesh, T, LO, adf_h, adf_v, hnn, **kw):
        h, v, fac = orig(mesh, T, LO, adf_h, adf_v, hnn, **kw)
        # apply no limiting at all (factor 1), as if plus/minus were mixed up to 1
        return adf_h, np.where(mesh.node_iface_mask, adf_v, 0.0), fac
    monkeypatch.setattr(O, "zalesak_limit", swapped)
    out, _, p = _step(m, c, c["T"][0], c["T_old"][0], return_parts=True)
    w = m.node_layer_mask
    up = p["LO"] + p["fct_ttf_max"]
    dn = p["LO"] + p["fct_ttf_min"]
    assert np.any(out[w] > up[w] + 1e-9) or np.any(out[w] < dn[w] - 1e-9)


def _ab2_consistent(T):
    """A two-level time extrapolation is consistent iff a field that does not change
    in time extrapolates to itself (the AB2 coefficients sum to one)."""
    ttf, told = O.tracer_ab(T, T)
    return told is T and np.max(np.abs(ttf - T)) <= 4e-16 * np.max(np.abs(T))


def test_ab2_extrapolation_consistent(pi_case):
    assert _ab2_consistent(pi_case["T"][0])


def test_planted_ab2_coefficient_breaks_consistency(pi_case, monkeypatch):
    monkeypatch.setattr(O, "_AB_NEW", 1.5)      # drops the +0.1 stabilisation
    assert not _ab2_consistent(pi_case["T"][0])


# --------------------------------------------------------------------------
# w-split implicit branch (a6)
# --------------------------------------------------------------------------
def test_vert_impl_identity_for_zero_w_and_conserves(pi_mesh, pi_case):
    m, c = pi_mesh, pi_case
    LO = c["T"][0]
    out = O.adv_tra_vert_impl(m, np.zeros_like(LO), LO, c["hnode_new"], dt=c["dt"])
    assert np.array_equal(out, LO)
    rng = np.random.default_rng(3)
    w_i = np.where(m.node_iface_mask, 1e-5 * rng.standard_normal(LO.shape), 0.0)
    w_i[:, 0] = 0.0                  # no flux through the surface: column mass conserved
    out = O.adv_tra_vert_impl(m, w_i, LO, c["hnode_new"], dt=c["dt"])
    col0 = np.sum(m.areasvol * c["hnode_new"] * LO * m.node_layer_mask, axis=1)
    col1 = np.sum(m.areasvol * c["hnode_new"] * out * m.node_layer_mask, axis=1)
    assert np.max(np.abs(col1 - col0) / np.abs(col0)) < 1e-12


# ---- the upwind-only driver (tracer_adv.py:154-165) ---------------------------------
def test_upwind_preserves_constant(icosa4_mesh):
    from fesom_jax_amd.inputs import make_case
    m = icosa4_mesh
    c = make_case(m, seed=81, dt=2700.0)
    T = np.where(m.node_layer_mask, 7.25, 0.0)
    Tn, Tr = O.advect_one(m, c["uv"], c["w_e"], c["helem"], c["hnode"], c["hnode_new"], T, T,
                          dt=c["dt"])
    assert Tr is T
    assert np.abs(Tn - T)[m.node_layer_mask].max() <= 1e-12 * 7.25


def test_upwind_is_affine_in_the_tracer(icosa4_mesh):
    """T_new is linear in (T, T_old) for a fixed flow: superposition holds to rounding."""
    from fesom_jax_amd.inputs import make_case
    m = icosa4_mesh
    c = make_case(m, seed=82, dt=2700.0)
    rng = np.random.default_rng(3)
    f = lambda T, To: O.advect_one(m, c["uv"], c["w_e"], c["helem"], c["hnode"],
                                   c["hnode_new"], T, To, dt=c["dt"])[0]
    a, b = c["T"][0], rng.standard_normal(c["T"][0].shape) * m.node_layer_mask
    ao, bo = 0.99 * a, rng.standard_normal(a.shape) * m.node_layer_mask
    lhs = f(a + 2.0 * b, ao + 2.0 * bo)
    rhs = f(a, ao) + 2.0 * f(b, bo)
    assert np.abs(lhs - rhs).max() <= 1e-11 * np.abs(lhs).max()


def test_upwind_mass_budget(pi_mesh, pi_case):
    """Sum of areasvol*hnode_new*T_new equals sum of areasvol*hnode*T plus dt times the
    surface vertical flux of ttfAB (the horizontal fluxes cancel between nodes)."""
    import math
    m, c = pi_mesh, pi_case
    T = c["T"][0]
    To = 0.99 * T
    Tn, _ = O.advect_one(m, c["uv"], c["w_e"], c["helem"], c["hnode"], c["hnode_new"], T, To,
                         dt=c["dt"])
    A, _ = O.tracer_ab(T, To)
    fv = O.adv_flux_ver(m, c["w_e"], A)
    lm = m.node_layer_mask
    before = math.fsum((m.areasvol * c["hnode"] * T)[lm])
    after = math.fsum((m.areasvol * c["hnode_new"] * Tn)[lm])
    surf = math.fsum(fv[:, 0][lm[:, 0]] * c["dt"])
    assert abs(after - before - surf) <= 1e-13 * abs(before)


def test_planted_upwind_mistake_breaks_constant(icosa4_mesh, monkeypatch):
    """Using T instead of ttfAB for the vertical flux (a dropped AB2 term) breaks the
    constant-preservation check when T_old differs from T."""
    from fesom_jax_amd.inputs import make_case
    m = icosa4_mesh
    c = make_case(m, seed=83, dt=2700.0)
    T = np.where(m.node_layer_mask, 3.0, 0.0)
    To = np.where(m.node_layer_mask, 2.0, 0.0)
    good = O.advect_one(m, c["uv"], c["w_e"], c["helem"], c["hnode"], c["hnode_new"], T, To,
                        dt=c["dt"])[0]
    # with T = 3 and T_old = 2 the AB2 field is the constant 3.6: still preserved exactly
    assert np.abs(good - T)[m.node_layer_mask].max() <= 1e-12 * 3.0
    orig = O.adv_flux_ver
    monkeypatch.setattr(O, "adv_flux_ver", lambda mesh, w, ttf: orig(mesh, w, ttf * 0.9))
    bad = O.advect_one(m, c["uv"], c["w_e"], c["helem"], c["hnode"], c["hnode_new"], T, To,
                       dt=c["dt"])[0]
    assert np.abs(bad - T)[m.node_layer_mask].max() > 1e-9
