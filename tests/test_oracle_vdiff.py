"""Pins for the implicit vertical diffusion oracle (oracle/vert_diff.py, restating
fesom_jax/tracer_diff.py:68-141).

The central check is written independently of the tridiagonal coefficients: the solve
must satisfy the backward-Euler flux form of vertical diffusion in every wet layer,

    areasvol_k hnode_new_k (T_new_k - T_k) = dt (q_k - q_{k+1}) + areasvol_k src_k,
    q_k = Kv_k area_k (T_new_{k-1} - T_new_k) / (Z_{k-1} - Z_k)   (interior interfaces),

with no flux through the surface and bottom interfaces (the reference's tridiagonal
comment, fesom_jax/tracer_diff.py:38-44). The reference's own property test
(fesom_jax/tests/test_tracers.py:214-235) adds a constant-in-z tracer staying put, the
column integral conserved and a gradient smoothed. Planted mistakes show that a swapped
spacing or a dropped sub-diagonal fails the flux form.
"""

import numpy as np
import pytest

from oracle import fct_numpy as O
from oracle import vert_diff as VD


def _kv(mesh, seed, lo=1e-5, hi=3e-2):
    rng = np.random.default_rng(seed)
    return np.where(mesh.node_iface_mask, rng.uniform(lo, hi, (mesh.nod2D, mesh.nl)), 0.0)


def _layer_depths(mesh, Z3d):
    if Z3d is not None:
        return np.asarray(Z3d)
    Z = np.asarray(mesh.Z)
    return np.broadcast_to(np.concatenate([Z, Z[-1:]]), (mesh.nod2D, mesh.nl))


def _flux_form_residual(mesh, T, Tn, Kv, hnn, dt, bc=None, sw=None, Z3d=None):
    """max |residual| / max |term| over wet layers of the flux form above."""
    Zl = _layer_depths(mesh, Z3d)
    av = np.asarray(mesh.areasvol)
    area = np.asarray(mesh.area)
    nzmin = np.asarray(mesh.ulevels_nod2D) - 1
    nzmax = np.asarray(mesh.nlevels_nod2D) - 1
    N, nl = T.shape
    # interface fluxes q_k between layers k-1 and k, zero at a column's top and bottom
    q = np.zeros((N, nl + 1))
    for k in range(1, nl):
        interior = (k > nzmin) & (k < nzmax)
        dz = Zl[:, k - 1] - Zl[:, k]
        q[:, k] = np.where(interior, Kv[:, k] * area[:, k] * (Tn[:, k - 1] - Tn[:, k])
                           / np.where(interior, dz, 1.0), 0.0)
    res, scale = 0.0, 0.0
    for k in range(nl):
        wet = (k >= nzmin) & (k < nzmax)
        lhs = av[:, k] * hnn[:, k] * (Tn[:, k] - T[:, k])
        rhs = dt * (q[:, k] - q[:, k + 1])
        if bc is not None:
            rhs = rhs + np.where(k == nzmin, av[:, k] * bc, 0.0)
        if sw is not None:
            below = sw[:, k + 1] * area[:, k + 1] if k + 1 < nl else 0.0
            rhs = rhs + dt * (av[:, k] * sw[:, k] - below)
        r = np.where(wet, lhs - rhs, 0.0)
        res = max(res, float(np.abs(r).max()))
        scale = max(scale, float(np.abs(np.where(wet, lhs, 0.0)).max()),
                    float(np.abs(np.where(wet, dt * q[:, k], 0.0)).max()))
    return res / max(scale, 1e-300)


@pytest.fixture(scope="module")
def vd_case(pi_mesh, pi_case):
    m = pi_mesh
    rng = np.random.default_rng(5)
    bc = rng.standard_normal(m.nod2D) * 0.3
    sw = np.where(m.node_iface_mask, 1e-4 * np.exp(np.asarray(m.zbar)[None, :] / 30.0), 0.0)
    sw = sw * (1.0 + 0.1 * rng.standard_normal(sw.shape))
    return dict(T=pi_case["T"][0], S=pi_case["T"][1], Kv=_kv(m, 6),
                hnn=pi_case["hnode_new"], dt=pi_case["dt"], bc=bc, sw=sw)


def test_solution_satisfies_flux_form(pi_mesh, vd_case):
    c = vd_case
    Tn = VD.impl_vert_diff_one(pi_mesh, c["T"], c["Kv"], c["hnn"], dt=c["dt"])
    assert np.abs(Tn - c["T"])[pi_mesh.node_layer_mask].max() > 1e-6   # diffusion acted
    assert _flux_form_residual(pi_mesh, c["T"], Tn, c["Kv"], c["hnn"], c["dt"]) < 1e-12


def test_flux_form_with_surface_increment_and_shortwave(pi_mesh, vd_case):
    c = vd_case
    Tn = VD.impl_vert_diff_one(pi_mesh, c["T"], c["Kv"], c["hnn"], dt=c["dt"],
                               bc_surf=c["bc"], sw_3d=c["sw"])
    r = _flux_form_residual(pi_mesh, c["T"], Tn, c["Kv"], c["hnn"], c["dt"], bc=c["bc"],
                            sw=c["sw"])
    assert r < 1e-12


def test_flux_form_with_live_depths(icosa4_mesh, icosa4_case):
    m, c = icosa4_mesh, icosa4_case
    rng = np.random.default_rng(8)
    Zs = np.concatenate([m.Z, m.Z[-1:]])
    Z3d = Zs[None, :] * (1.0 + 0.05 * rng.uniform(-1, 1, (m.nod2D, 1)))
    Kv = _kv(m, 9)
    Tn = VD.impl_vert_diff_one(m, c["T"][0], Kv, c["hnode_new"], dt=c["dt"], Z3d=Z3d)
    assert _flux_form_residual(m, c["T"][0], Tn, Kv, c["hnode_new"], c["dt"], Z3d=Z3d) < 1e-12


def test_reference_properties(pi_mesh, vd_case):
    """fesom_jax/tests/test_tracers.py:214-235: Kv = 1 on every interface."""
    m, c = pi_mesh, vd_case
    wet = m.node_layer_mask
    Kv_big = np.where(m.node_iface_mask, 1.0, 0.0)
    Tc = np.where(wet, 20.0, 0.0)
    Tc2 = VD.impl_vert_diff_one(m, Tc, Kv_big, c["hnn"], dt=c["dt"])
    assert np.max(np.abs(Tc2[wet] - 20.0)) < 1e-12
    k = np.arange(m.nl)[None, :]
    Tg = np.where(wet, 10.0 + 0.5 * k, 0.0)
    Tg2 = VD.impl_vert_diff_one(m, Tg, Kv_big, c["hnn"], dt=c["dt"])
    av, hn = np.asarray(m.areasvol), np.asarray(c["hnn"])
    i0 = np.sum(av * hn * Tg, axis=1)
    i1 = np.sum(av * hn * Tg2, axis=1)
    assert np.max(np.abs(i1 - i0) / np.maximum(np.abs(i0), 1e-30)) < 1e-12
    rng0 = np.where(wet, Tg, -1e30).max(1) - np.where(wet, Tg, 1e30).min(1)
    rng1 = np.where(wet, Tg2, -1e30).max(1) - np.where(wet, Tg2, 1e30).min(1)
    assert rng1.mean() < rng0.mean()


def test_zero_diffusivity_is_identity(pi_mesh, vd_case):
    c = vd_case
    Tn = VD.impl_vert_diff_one(pi_mesh, c["T"], np.zeros_like(c["Kv"]), c["hnn"], dt=c["dt"])
    assert np.array_equal(Tn, c["T"])


def test_surface_increment_closed_form(pi_mesh, vd_case):
    """Kv = 0: only the surface layer moves, by exactly bc / hnode_new."""
    m, c = pi_mesh, vd_case
    Tn = VD.impl_vert_diff_one(m, c["T"], np.zeros_like(c["Kv"]), c["hnn"], dt=c["dt"],
                               bc_surf=c["bc"])
    wet = m.node_layer_mask
    top = np.asarray(m.ulevels_nod2D) - 1
    rows = np.arange(m.nod2D)
    want = c["T"].copy()
    has = wet[rows, top]
    want[rows[has], top[has]] = c["T"][rows[has], top[has]] + c["bc"][has] / c["hnn"][rows[has], top[has]]
    assert np.array_equal(Tn, want)


def test_shortwave_heat_budget(pi_mesh, vd_case):
    """With diffusion on, the column heat content changes by the shortwave flux through
    the top of the column minus what leaves through the bottom interface."""
    m, c = pi_mesh, vd_case
    Tn = VD.impl_vert_diff_one(m, c["T"], c["Kv"], c["hnn"], dt=c["dt"], sw_3d=c["sw"])
    av, area, sw = np.asarray(m.areasvol), np.asarray(m.area), c["sw"]
    wet = m.node_layer_mask
    d = np.sum(np.where(wet, av * c["hnn"] * (Tn - c["T"]), 0.0), axis=1)
    top = np.asarray(m.ulevels_nod2D) - 1
    bot = np.asarray(m.nlevels_nod2D) - 1
    rows = np.arange(m.nod2D)
    want = c["dt"] * (av[rows, top] * sw[rows, top] - area[rows, bot] * sw[rows, bot])
    # with av == area at the top interface this is the classic flux difference
    want = want + c["dt"] * np.sum(np.where(wet[:, 1:], (av[:, 1:] - area[:, 1:]) * sw[:, 1:], 0.0),
                                   axis=1)
    wet_cols = wet.any(axis=1)
    scale = np.abs(want[wet_cols]).max()
    assert np.abs(d - want)[wet_cols].max() <= 1e-11 * scale


def test_static_and_equal_live_depths_agree_bitwise(pi_mesh, vd_case):
    m, c = pi_mesh, vd_case
    Zs = np.concatenate([m.Z, m.Z[-1:]])
    Z3d = np.broadcast_to(Zs, (m.nod2D, m.nl)).copy()
    a = VD.impl_vert_diff_one(m, c["T"], c["Kv"], c["hnn"], dt=c["dt"])
    b = VD.impl_vert_diff_one(m, c["T"], c["Kv"], c["hnn"], dt=c["dt"], Z3d=Z3d)
    assert np.array_equal(a, b)


def test_two_tracer_form_and_floor(pi_mesh, vd_case):
    m, c = pi_mesh, vd_case
    T2, S2 = VD.impl_vert_diff(m, c["T"], c["S"], c["Kv"], c["hnn"], dt=c["dt"],
                               bc_T=c["bc"], sw_3d=c["sw"])
    assert np.array_equal(T2, VD.impl_vert_diff_one(m, c["T"], c["Kv"], c["hnn"], dt=c["dt"],
                                                    bc_surf=c["bc"], sw_3d=c["sw"]))
    assert np.array_equal(S2, VD.impl_vert_diff_one(m, c["S"], c["Kv"], c["hnn"], dt=c["dt"]))
    low = np.where(m.node_layer_mask, 0.1, -3.0)
    f = VD.salinity_floor(m, low)
    assert np.all(f[m.node_layer_mask] == VD.S_FLOOR) and np.all(f[~m.node_layer_mask] == -3.0)


# ---- planted mistakes: each must fail the flux form ----------------------------------
def test_planted_swapped_spacings_fail(pi_mesh, vd_case, monkeypatch):
    c = vd_case
    real = VD._spacings
    monkeypatch.setattr(VD, "_spacings", lambda mesh, Z3d: real(mesh, Z3d)[::-1])
    Tn = VD.impl_vert_diff_one(pi_mesh, c["T"], c["Kv"], c["hnn"], dt=c["dt"])
    assert _flux_form_residual(pi_mesh, c["T"], Tn, c["Kv"], c["hnn"], c["dt"]) > 1e-6


def test_planted_dropped_subdiagonal_fails(pi_mesh, vd_case, monkeypatch):
    c = vd_case
    monkeypatch.setattr(VD, "tdma", lambda a, b, cc, d: O.tdma(np.zeros_like(a), b, cc, d))
    Tn = VD.impl_vert_diff_one(pi_mesh, c["T"], c["Kv"], c["hnn"], dt=c["dt"])
    assert _flux_form_residual(pi_mesh, c["T"], Tn, c["Kv"], c["hnn"], c["dt"]) > 1e-6
