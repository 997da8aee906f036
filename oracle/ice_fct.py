"""CPU oracle (test infrastructure only): a numpy float64 restatement of the reference's
sea-ice FCT advection, fesom_jax/ice_adv.py:37-194 (fct_solve and its parts): the
Taylor-Galerkin right-hand side, the consistent-mass low- and high-order solves and the
element-centric Zalesak limiter, for the three ice tracers on the surface nodes.

Only ``tests/`` uses this module, and only as the checker; the product path
(``fesom_jax_amd``) never imports it.

Operation order follows the reference's expressions left to right, no fused
multiply-add; sums over an element's three vertices run ((x0 + x1) + x2); node sums add
element contributions in increasing element id from +0.0 (``scatter_add``), the order
the GPU kernels keep. Pinned in tests/test_oracle_ice.py.
"""

from __future__ import annotations

import numpy as np

from .fct_numpy import scatter_add

ICE_SCALE_AREA = 2.0e8          # fesom_jax/ice_adv.py:31
ICE_FLUX_EPS = 1.0e-12          # fesom_jax/ice_adv.py:32


class IceConfig:
    """The three settings the ice FCT reads (fesom_jax/ice.py:58-62 defaults)."""

    def __init__(self, ice_dt=500.0, ice_diff=10.0, ice_gamma_fct=0.5):
        self.ice_dt = ice_dt
        self.ice_diff = ice_diff
        self.ice_gamma_fct = ice_gamma_fct


def _sum3(x):
    return (x[:, 0] + x[:, 1]) + x[:, 2]


def mm_times(mesh, X):
    """Consistent-mass product (fesom_jax/ice_adv.py:37-44)."""
    en = np.asarray(mesh.elem_nodes)
    Xv = np.asarray(X, np.float64)[en]
    contrib = (np.asarray(mesh.elem_area) / 12.0)[:, None] * (Xv + _sum3(Xv)[:, None])
    contrib = np.where((np.asarray(mesh.ulevels) <= 1)[:, None], contrib, 0.0)
    return scatter_add(contrib, en, mesh.nod2D)


def tg_entries(cfg, mesh, u_ice, v_ice):
    """The element matrices vol*dt*(a - b - c) of the Taylor-Galerkin RHS
    (fesom_jax/ice_adv.py:47-73): [elem, n, q]."""
    en = np.asarray(mesh.elem_nodes)
    gs = np.asarray(mesh.gradient_sca)
    dx, dy = gs[:, 0:3], gs[:, 3:6]
    vol = np.asarray(mesh.elem_area)
    dt = cfg.ice_dt
    U = np.asarray(u_ice, np.float64)[en]
    V = np.asarray(v_ice, np.float64)[en]
    um, vm = _sum3(U), _sum3(V)
    diff = cfg.ice_diff * np.sqrt(vol / ICE_SCALE_AREA)
    a = (dx[:, :, None] * (um[:, None, None] + U[:, None, :])
         + dy[:, :, None] * (vm[:, None, None] + V[:, None, :])) / 12.0
    b = diff[:, None, None] * (dx[:, :, None] * dx[:, None, :] + dy[:, :, None] * dy[:, None, :])
    udn = um[:, None] * dx + vm[:, None] * dy
    c = 0.5 * dt * (udn[:, :, None] * udn[:, None, :]) / 9.0
    return vol[:, None, None] * dt * (a - b - c)


def tg_rhs(cfg, mesh, entries, X):
    en = np.asarray(mesh.elem_nodes)
    Xv = np.asarray(X, np.float64)[en]
    p = entries * Xv[:, None, :]
    contrib = (p[:, :, 0] + p[:, :, 1]) + p[:, :, 2]
    contrib = np.where((np.asarray(mesh.ulevels) <= 1)[:, None], contrib, 0.0)
    return scatter_add(contrib, en, mesh.nod2D)


def _inv_area(mesh):
    area = np.asarray(mesh.area)[:, 0]
    return 1.0 / np.where(area > 0, area, 1.0)


def solve_low_order(cfg, mesh, X, rhs):
    """fesom_jax/ice_adv.py:76-87."""
    g = cfg.ice_gamma_fct
    nc = np.asarray(mesh.ulevels_nod2D) <= 1
    X = np.asarray(X, np.float64)
    return np.where(nc, (rhs + g * mm_times(mesh, X)) * _inv_area(mesh) + (1.0 - g) * X, X)


def solve_high_order(cfg, mesh, rhs):
    """fesom_jax/ice_adv.py:90-108 (two residual corrections)."""
    inv_area = _inv_area(mesh)
    nc = np.asarray(mesh.ulevels_nod2D) <= 1
    d = np.where(nc, rhs * inv_area, 0.0)
    for _ in range(2):
        d = np.where(nc, d + (rhs - mm_times(mesh, d)) * inv_area, d)
    return d


def fem_fct(cfg, mesh, vals, vals_l, dvals, *, return_parts=False):
    """Zalesak limiter of one tracer (fesom_jax/ice_adv.py:111-174)."""
    en = np.asarray(mesh.elem_nodes)
    g = cfg.ice_gamma_fct
    inv_area = _inv_area(mesh)
    nc_e = np.asarray(mesh.ulevels) <= 1
    nc_n = np.asarray(mesh.ulevels_nod2D) <= 1
    vol = np.asarray(mesh.elem_area)
    vals = np.asarray(vals, np.float64)
    w = g * vals + dvals
    wv = w[en]
    s = _sum3(wv)[:, None] - 3.0 * wv
    fluxes = np.where(nc_e[:, None], -s * (vol[:, None] * inv_area[en]) / 12.0, 0.0)
    vl, vv = vals_l[en], vals[en]
    elem_hi = np.maximum(vl.max(1), vv.max(1))
    elem_lo = np.minimum(vl.min(1), vv.min(1))
    hi = np.full(mesh.nod2D, -np.inf)
    lo = np.full(mesh.nod2D, np.inf)
    for q in range(3):
        np.maximum.at(hi, en[:, q], np.where(nc_e, elem_hi, -np.inf))
        np.minimum.at(lo, en[:, q], np.where(nc_e, elem_lo, np.inf))
    hi = np.where(np.isfinite(hi), hi, 0.0)
    lo = np.where(np.isfinite(lo), lo, 0.0)
    tmax = hi - vals_l
    tmin = lo - vals_l
    fp = scatter_add(np.maximum(fluxes, 0.0), en, mesh.nod2D)
    fm = scatter_add(np.minimum(fluxes, 0.0), en, mesh.nod2D)
    with np.errstate(divide="ignore", invalid="ignore"):
        pplus = np.where(fp > 0.0,
                         np.minimum(tmax / np.where(fp > ICE_FLUX_EPS, fp, ICE_FLUX_EPS), 1.0), 0.0)
        pminus = np.where(fm < 0.0,
                          np.minimum(tmin / np.where(fm < -ICE_FLUX_EPS, fm, -ICE_FLUX_EPS), 1.0),
                          0.0)
    pplus = np.where(nc_n, pplus, 0.0)
    pminus = np.where(nc_n, pminus, 0.0)
    cand = np.where(fluxes >= 0.0, pplus[en], pminus[en])
    ae = cand.min(axis=1)
    limited = np.where(nc_e[:, None], fluxes * ae[:, None], 0.0)
    out = vals_l + scatter_add(limited, en, mesh.nod2D)
    if return_parts:
        return out, dict(hi=hi, lo=lo, fluxes=fluxes, pplus=pplus, pminus=pminus, ae=ae)
    return out


def fct_solve(cfg, mesh, a_ice, m_ice, m_snow, u_ice, v_ice):
    """One FCT step of the three ice tracers (fesom_jax/ice_adv.py:177-194); returns
    ``(a_new, m_new, ms_new)``."""
    E = tg_entries(cfg, mesh, u_ice, v_ice)
    rhs_m, rhs_a, rhs_ms = tg_rhs(cfg, mesh, E, m_ice), tg_rhs(cfg, mesh, E, a_ice), \
        tg_rhs(cfg, mesh, E, m_snow)
    da, dm, dms = (solve_high_order(cfg, mesh, r) for r in (rhs_a, rhs_m, rhs_ms))
    a_l = solve_low_order(cfg, mesh, a_ice, rhs_a)
    m_l = solve_low_order(cfg, mesh, m_ice, rhs_m)
    ms_l = solve_low_order(cfg, mesh, m_snow, rhs_ms)
    m_new = fem_fct(cfg, mesh, m_ice, m_l, dm)
    a_new = fem_fct(cfg, mesh, a_ice, a_l, da)
    ms_new = fem_fct(cfg, mesh, m_snow, ms_l, dms)
    return a_new, m_new, ms_new
