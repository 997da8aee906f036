"""CPU oracle (test infrastructure only): a numpy float64 restatement of the reference's
FCT tracer advection, fesom_jax/tracer_adv.py:43-578, with its primitives
fesom_jax/ops.py:56-168.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
use this module, and only as the checker; the product path (``fesom_jax_amd``) never
imports it.

Every stage keeps the reference's operation order (left-to-right Python evaluation of
the same expressions, no fused multiply-add) so that a GPU kernel compiled without FP
contraction and summing each node's contributions in the same order reproduces it
bit for bit. Segment sums add each segment's contributions sequentially in increasing
flattened-index order starting from +0.0, which is what ``np.add.at`` does.

Pinning (tests/test_oracle_*.py): the loop restatement in ``oracle/fct_loops.c``
(independent code following the C-kernel semantics restated by the reference's own
``_fct_ref``, fesom_jax/tests/test_tracers.py:325-516); constant-tracer preservation
(fesom_jax/tests/test_tracers.py:599-606); closed mass budget; Zalesak bounds;
linearity when the limiter is inactive; exact gradient of linear fields; TDMA against
``np.linalg.solve`` (fesom_jax/tests/test_ops.py:161-176); scatter-add against a
brute-force loop (fesom_jax/tests/test_ops.py:45-106). Each test names the planted
mistake it would catch.
"""

from __future__ import annotations

import numpy as np

R_EARTH = 6367500.0                       # fesom_jax/config.py:41
DT_DEFAULT = 2400.0                       # fesom_jax/config.py:100
_EPS_AB = 0.1                             # fesom_jax/tracer_adv.py:44
_AB_OLD = -(0.5 + _EPS_AB)                # fesom_jax/tracer_adv.py:45
_AB_NEW = (1.5 + _EPS_AB)                 # fesom_jax/tracer_adv.py:46
_FCT_FLUX_EPS = 1e-16                     # fesom_jax/tracer_adv.py:192
_FCT_BIG = 1e3                            # fesom_jax/tracer_adv.py:193
# Coefficients of the two high-order reconstructions, named so that the pinning tests
# can plant a wrong value (tests/test_oracle_pins.py): MFCT edge means
# ``T1 + (2*diff + ...)/6`` (fesom_jax/tracer_adv.py:222-223) and QR4C interface means
# ``T + (2*qc + qu)*(zbar - Z)/3`` (fesom_jax/tracer_adv.py:282-283).
_MFCT_DIFF_W = 2.0
_MFCT_DEN = 6.0
_QR4C_QC_W = 2.0
_QR4C_DEN = 3.0


# --------------------------------------------------------------------------
# primitives (fesom_jax/ops.py)
# --------------------------------------------------------------------------
def _segment_slots(seg_flat: np.ndarray, num_segments: int):
    """Group flat indices by segment, keeping increasing index order inside a
    segment. Returns a list of (segment_ids, flat_indices) pairs, one per slot
    position: slot j holds the j-th contribution of every segment that has one."""
    valid = seg_flat >= 0
    idx = np.nonzero(valid)[0]
    s = seg_flat[idx]
    order = np.argsort(s, kind="stable")
    s_sorted = s[order]
    idx_sorted = idx[order]
    counts = np.bincount(s_sorted, minlength=num_segments)
    starts = np.zeros(num_segments + 1, np.int64)
    starts[1:] = np.cumsum(counts)
    pos = np.arange(s_sorted.size) - starts[s_sorted]
    slots = []
    for j in range(int(counts.max()) if counts.size and s_sorted.size else 0):
        sel = pos == j
        slots.append((s_sorted[sel], idx_sorted[sel]))
    return slots


def scatter_add(vals, seg, num_segments: int):
    """Masked segment sum (fesom_jax/ops.py:56-81): ``seg < 0`` contributes nothing;
    each segment accumulates in increasing flat-index order from +0.0."""
    vals = np.asarray(vals)
    seg = np.asarray(seg)
    lead = seg.ndim
    if vals.shape[:lead] != seg.shape:
        raise ValueError(f"scatter_add: vals leading dims {vals.shape[:lead]} must match "
                         f"seg shape {seg.shape}")
    seg_flat = seg.reshape(-1).astype(np.int64)
    vals_flat = vals.reshape((seg_flat.shape[0],) + vals.shape[lead:])
    out = np.zeros((num_segments,) + vals.shape[lead:], dtype=np.float64)
    for segs, idx in _segment_slots(seg_flat, num_segments):
        out[segs] = out[segs] + vals_flat[idx]
    return out


def segment_max(vals, seg_flat, num_segments):
    """``jax.ops.segment_max`` semantics: -inf for an empty segment."""
    out = np.full((num_segments,) + vals.shape[1:], -np.inf)
    for segs, idx in _segment_slots(np.asarray(seg_flat, np.int64), num_segments):
        out[segs] = np.maximum(out[segs], vals[idx])
    return out


def segment_min(vals, seg_flat, num_segments):
    """``jax.ops.segment_min`` semantics: +inf for an empty segment."""
    out = np.full((num_segments,) + vals.shape[1:], np.inf)
    for segs, idx in _segment_slots(np.asarray(seg_flat, np.int64), num_segments):
        out[segs] = np.minimum(out[segs], vals[idx])
    return out


def mask_below_bottom(field, level_mask):
    """fesom_jax/ops.py:105-115."""
    field = np.asarray(field)
    m = np.asarray(level_mask)
    if field.ndim > m.ndim:
        m = m.reshape(m.shape + (1,) * (field.ndim - m.ndim))
    return np.where(m, field, 0.0)


def tdma(a, b, c, d):
    """Thomas algorithm along the last axis (fesom_jax/ops.py:121-168): forward sweep
    ``m = b - a*cp_prev; cp = c/m; dp = (d - a*dp_prev)/m`` then
    ``x = dp - cp*x_next``."""
    a, b, c, d = (np.asarray(v, np.float64) for v in (a, b, c, d))
    K = b.shape[-1]
    shape = np.broadcast_shapes(a.shape, b.shape, c.shape, d.shape)
    a, b, c, d = (np.broadcast_to(v, shape) for v in (a, b, c, d))
    cp = np.zeros(shape)
    dp = np.zeros(shape)
    cp_prev = np.zeros(shape[:-1])
    dp_prev = np.zeros(shape[:-1])
    for i in range(K):
        m = b[..., i] - a[..., i] * cp_prev
        cp[..., i] = c[..., i] / m
        dp[..., i] = (d[..., i] - a[..., i] * dp_prev) / m
        cp_prev, dp_prev = cp[..., i], dp[..., i]
    x = np.zeros(shape)
    x_next = np.zeros(shape[:-1])
    for i in range(K - 1, -1, -1):
        x[..., i] = dp[..., i] - cp[..., i] * x_next
        x_next = x[..., i]
    return x


def _shift_down(x):
    """``out[..., k] = x[..., k-1]``, replicated at k=0 (fesom_jax/tracer_adv.py:49-51)."""
    return np.concatenate([x[..., :1], x[..., :-1]], axis=-1)


def _shift_up(x):
    """``out[..., k] = x[..., k+1]``, zero at the last level (fesom_jax/tracer_adv.py:54-56)."""
    return np.concatenate([x[..., 1:], np.zeros_like(x[..., :1])], axis=-1)


def _safe(x):
    return np.where(x > 0.0, x, 1.0)


# --------------------------------------------------------------------------
# stages (fesom_jax/tracer_adv.py)
# --------------------------------------------------------------------------
def tracer_ab(values, valuesold):
    """fesom_jax/tracer_adv.py:59-66."""
    ttfAB = _AB_OLD * valuesold + _AB_NEW * values
    return ttfAB, values


def edge_vflux(mesh, uv, helem):
    """fesom_jax/tracer_adv.py:69-93."""
    et = np.asarray(mesh.edge_tri)
    el1, el2 = et[:, 0], et[:, 1]
    has1, has2 = el1 >= 0, el2 >= 0
    el1s = np.where(has1, el1, 0)
    el2s = np.where(has2, el2, 0)
    cross = np.asarray(mesh.edge_cross_dxdy)
    dx1, dy1 = cross[:, 0:1], cross[:, 1:2]
    dx2, dy2 = cross[:, 2:3], cross[:, 3:4]
    U, V = uv[:, :, 0], uv[:, :, 1]
    lm = np.asarray(mesh.elem_layer_mask)
    t1 = np.where(lm[el1s] & has1[:, None], (U[el1s] * dy1 - V[el1s] * dx1) * helem[el1s], 0.0)
    t2 = np.where(lm[el2s] & has2[:, None], (V[el2s] * dx2 - U[el2s] * dy2) * helem[el2s], 0.0)
    return t1 + t2


def adv_flux_hor(mesh, uv, helem, ttf):
    """fesom_jax/tracer_adv.py:96-105."""
    vflux = edge_vflux(mesh, uv, helem)
    T1 = ttf[mesh.edges[:, 0]]
    T2 = ttf[mesh.edges[:, 1]]
    av = np.abs(vflux)
    return -0.5 * (T1 * (vflux + av) + T2 * (vflux - av))


def adv_flux_ver(mesh, w_e, ttf):
    """fesom_jax/tracer_adv.py:108-120."""
    w = w_e
    aw = np.abs(w)
    T_below = ttf
    T_above = _shift_down(ttf)
    flux = -0.5 * (T_below * (w + aw) + T_above * (w - aw)) * mesh.area
    return mask_below_bottom(flux, mesh.node_layer_mask)


def flux2dtracer(mesh, flux_h, flux_v, *, dt=DT_DEFAULT):
    """fesom_jax/tracer_adv.py:123-139 (upwind-only driver's assembly)."""
    areasvol = mesh.areasvol
    safe = np.where(areasvol > 0.0, areasvol, 1.0)
    lm = mesh.node_layer_mask
    dttf_v = np.where(lm, (flux_v - _shift_up(flux_v)) * dt / safe, 0.0)
    vals = np.stack([flux_h, -flux_h], axis=1)
    raw = scatter_add(vals, mesh.edges, mesh.nod2D)
    dttf_h = np.where(lm, raw * dt / safe, 0.0)
    return dttf_h + dttf_v


def ale_reconstruct(mesh, T, del_ttf, hnode, hnode_new):
    """fesom_jax/tracer_adv.py:142-151."""
    del_ttf = del_ttf + T * (hnode - hnode_new)
    safe_hn = np.where(hnode_new > 0.0, hnode_new, 1.0)
    upd = mesh.node_layer_mask & (hnode_new > 0.0)
    dT = np.where(upd, del_ttf / safe_hn, 0.0)
    return T + dT


def advect_one(mesh, uv, w_e, helem, hnode, hnode_new, T, T_old, *, dt=DT_DEFAULT):
    """Upwind-only driver, fesom_jax/tracer_adv.py:154-165."""
    ttfAB, T_old_new = tracer_ab(T, T_old)
    flux_h = adv_flux_hor(mesh, uv, helem, ttfAB)
    flux_v = adv_flux_ver(mesh, w_e, ttfAB)
    del_ttf = flux2dtracer(mesh, flux_h, flux_v, dt=dt)
    return ale_reconstruct(mesh, T, del_ttf, hnode, hnode_new), T_old_new


def adv_flux_hor_ho(mesh, uv, helem, ttf, eud):
    """MFCT high-order horizontal flux, fesom_jax/tracer_adv.py:196-225."""
    vflux = edge_vflux(mesh, uv, helem)
    n1, n2 = mesh.edges[:, 0], mesh.edges[:, 1]
    T1, T2 = ttf[n1], ttf[n2]
    diff = T2 - T1
    g0, g1, g2, g3 = eud[:, :, 0], eud[:, :, 1], eud[:, :, 2], eud[:, :, 3]
    et = mesh.edge_tri
    el1, el2 = et[:, 0], et[:, 1]
    has2 = el2 >= 0
    el1s = np.where(el1 >= 0, el1, 0)
    el2s = np.where(has2, el2, 0)
    a = R_EARTH * mesh.elem_cos[el1s]
    a = np.where(has2, 0.5 * (a + R_EARTH * mesh.elem_cos[el2s]), a)
    a = a[:, None]
    edx = mesh.edge_dxdy[:, 0:1]
    edy = mesh.edge_dxdy[:, 1:2]
    Tmean1 = T1 + (_MFCT_DIFF_W * diff + edx * a * g0 + edy * R_EARTH * g2) / _MFCT_DEN
    Tmean2 = T2 - (_MFCT_DIFF_W * diff + edx * a * g1 + edy * R_EARTH * g3) / _MFCT_DEN
    av = np.abs(vflux)
    return -0.5 * ((vflux + av) * Tmean1 + (vflux - av) * Tmean2)


def _z_stencil(Z, nl):
    """fesom_jax/tracer_adv.py:228-238."""
    Zp = np.concatenate([Z, Z[-1:]])
    Z_nz = Zp
    Z_m1 = np.concatenate([Zp[:1], Zp[:-1]])
    Z_m2 = np.concatenate([Zp[:2], Zp[:-2]])
    Z_p1 = np.concatenate([Zp[1:], Zp[-1:]])
    return Z_nz, Z_m1, Z_m2, Z_p1


def adv_flux_ver_ho(mesh, w_e, ttf, Z3d=None, zbar3=None):
    """QR4C high-order vertical flux, fesom_jax/tracer_adv.py:241-295."""
    W = w_e
    area = mesh.area
    T = ttf
    Tm1 = _shift_down(T)
    Tm2 = _shift_down(Tm1)
    Tp1 = _shift_up(T)
    surf = -T * W * area
    cent = -0.5 * (Tm1 + T) * W * area
    if Z3d is None:
        Z_nz, Z_m1, Z_m2, Z_p1 = _z_stencil(np.asarray(mesh.Z), mesh.nl)
        zb = np.asarray(mesh.zbar)
    else:
        Z_nz = Z3d
        Z_m1 = _shift_down(Z3d)
        Z_m2 = _shift_down(Z_m1)
        Z_p1 = np.concatenate([Z3d[:, 1:], Z3d[:, -1:]], axis=1)
        zb = zbar3

    def _sf(x):
        return np.where(x == 0.0, 1.0, x)

    with np.errstate(divide="ignore", invalid="ignore"):
        qc = (Tm1 - T) / _sf(Z_m1 - Z_nz)
        qu = (T - Tp1) / _sf(Z_nz - Z_p1)
        qd = (Tm2 - Tm1) / _sf(Z_m2 - Z_m1)
        Tmean1 = T + (_QR4C_QC_W * qc + qu) * (zb - Z_nz) / _QR4C_DEN
        Tmean2 = Tm1 + (_QR4C_QC_W * qc + qd) * (zb - Z_m1) / _QR4C_DEN
        quad = -0.5 * (Tmean1 + Tmean2) * W * area
    k = np.arange(mesh.nl)[None, :]
    nzmin = (mesh.ulevels_nod2D.astype(np.int64) - 1)[:, None]
    nzmax = (mesh.nlevels_nod2D.astype(np.int64) - 1)[:, None]
    is_surf = k == nzmin
    is_bot = k == nzmax
    is_cent = (k == nzmin + 1) | (k == nzmax - 1)
    is_int = (k >= nzmin + 2) & (k <= nzmax - 2)
    hi = np.where(is_surf, surf,
                  np.where(is_bot, 0.0,
                           np.where(is_cent, cent,
                                    np.where(is_int, quad, 0.0))))
    return mask_below_bottom(hi, mesh.node_iface_mask)


def compute_fct_lo(mesh, flux_h, flux_v, T, hnode, hnode_new, *, dt=DT_DEFAULT):
    """fesom_jax/tracer_adv.py:298-313."""
    vals = np.stack([flux_h, -flux_h], axis=1)
    div_h = scatter_add(vals, mesh.edges, mesh.nod2D)
    div_v = flux_v - _shift_up(flux_v)
    areasvol = mesh.areasvol
    safe_a = np.where(areasvol > 0.0, areasvol, 1.0)
    safe_hn = np.where(hnode_new > 0.0, hnode_new, 1.0)
    numer = T * hnode + (div_h + div_v) * dt / safe_a
    lo = numer / safe_hn
    valid = mesh.node_layer_mask & (hnode_new > 0.0) & (areasvol > 0.0)
    return np.where(valid, lo, 0.0)


def tracer_gradient_elements(mesh, ttf):
    """fesom_jax/tracer_adv.py:316-326."""
    t = ttf[mesh.elem_nodes]
    gs = mesh.gradient_sca
    gx = gs[:, 0:1] * t[:, 0] + gs[:, 1:2] * t[:, 1] + gs[:, 2:3] * t[:, 2]
    gy = gs[:, 3:4] * t[:, 0] + gs[:, 4:5] * t[:, 1] + gs[:, 5:6] * t[:, 2]
    m = mesh.elem_layer_mask
    return np.stack([np.where(m, gx, 0.0), np.where(m, gy, 0.0)], axis=-1)


def node_avg_grad(mesh, tr_xy):
    """fesom_jax/tracer_adv.py:329-341."""
    aw = np.where(mesh.elem_layer_mask, mesh.elem_area[:, None], 0.0)
    contrib = tr_xy * aw[:, :, None]
    en = mesh.elem_nodes
    e3 = (mesh.elem2D, 3)
    cb = np.broadcast_to(contrib[:, None], e3 + tr_xy.shape[1:])
    wb = np.broadcast_to(aw[:, None], e3 + (mesh.nl,))
    num = scatter_add(cb, en, mesh.nod2D)
    den = scatter_add(wb, en, mesh.nod2D)
    return num / np.where(den > 0.0, den, 1.0)[:, :, None]


def fill_up_dn_grad(mesh, tr_xy):
    """fesom_jax/tracer_adv.py:344-368."""
    gnode = node_avg_grad(mesh, tr_xy)
    up, dn = mesh.edge_up_dn_tri[:, 0], mesh.edge_up_dn_tri[:, 1]
    interior = (up >= 0) & (dn >= 0)
    gup = tr_xy[np.where(up >= 0, up, 0)]
    gdn = tr_xy[np.where(dn >= 0, dn, 0)]
    n1, n2 = mesh.edges[:, 0], mesh.edges[:, 1]
    gn1, gn2 = gnode[n1], gnode[n2]
    k = np.arange(mesh.nl)[None, :]
    nzmin = np.maximum(mesh.ulevels_nod2D_max[n1], mesh.ulevels_nod2D_max[n2]).astype(np.int64)[:, None] - 1
    nzmax = np.minimum(mesh.nlevels_nod2D_min[n1], mesh.nlevels_nod2D_min[n2]).astype(np.int64)[:, None] - 1
    use_tri = interior[:, None] & (k >= nzmin) & (k < nzmax)
    g0 = np.where(use_tri, gup[:, :, 0], gn1[:, :, 0])
    g1 = np.where(use_tri, gdn[:, :, 0], gn2[:, :, 0])
    g2 = np.where(use_tri, gup[:, :, 1], gn1[:, :, 1])
    g3 = np.where(use_tri, gdn[:, :, 1], gn2[:, :, 1])
    return np.stack([g0, g1, g2, g3], axis=-1)


def zalesak_limit(mesh, T, LO, adf_h, adf_v, hnode_new, *, dt=DT_DEFAULT, exch=None,
                  return_factors=False):
    """fesom_jax/tracer_adv.py:371-466. With ``return_factors`` also returns
    ``(plus_fac, minus_fac, fct_ttf_max, fct_ttf_min, fct_plus, fct_minus)``."""
    _exch = exch if exch is not None else (lambda f, kind: f)
    nl = mesh.nl
    nlm = mesh.node_layer_mask
    ttf_max = np.maximum(LO, T)
    ttf_min = np.minimum(LO, T)
    em = mesh.elem_layer_mask
    aux_max = np.where(em, np.max(ttf_max[mesh.elem_nodes], axis=1), -_FCT_BIG)
    aux_min = np.where(em, np.min(ttf_min[mesh.elem_nodes], axis=1), _FCT_BIG)
    seg = mesh.elem_nodes.reshape(-1)
    e3 = mesh.elem2D * 3
    cmax = np.broadcast_to(aux_max[:, None], (mesh.elem2D, 3, nl)).reshape(e3, nl)
    cmin = np.broadcast_to(aux_min[:, None], (mesh.elem2D, 3, nl)).reshape(e3, nl)
    tvert_max = segment_max(cmax, seg, mesh.nod2D)
    tvert_min = segment_min(cmin, seg, mesh.nod2D)
    k = np.arange(nl)[None, :]
    nzmin = (mesh.ulevels_nod2D.astype(np.int64) - 1)[:, None]
    nzmax = (mesh.nlevels_nod2D.astype(np.int64) - 1)[:, None]
    use_own = (k == nzmin) | (k == nzmax - 1)
    tmax3 = np.maximum(np.maximum(_shift_down(tvert_max), tvert_max), _shift_up(tvert_max))
    tmin3 = np.minimum(np.minimum(_shift_down(tvert_min), tvert_min), _shift_up(tvert_min))
    fct_ttf_max = np.where(use_own, tvert_max, tmax3) - LO
    fct_ttf_min = np.where(use_own, tvert_min, tmin3) - LO
    fct_ttf_max = np.where(nlm, fct_ttf_max, 0.0)
    fct_ttf_min = np.where(nlm, fct_ttf_min, 0.0)

    fv_bot = _shift_up(adf_v)
    pos_v = np.maximum(adf_v, 0.0) + np.maximum(-fv_bot, 0.0)
    neg_v = np.minimum(adf_v, 0.0) + np.minimum(-fv_bot, 0.0)
    pos_v = np.where(nlm, pos_v, 0.0)
    neg_v = np.where(nlm, neg_v, 0.0)
    ph = np.stack([np.maximum(adf_h, 0.0), np.maximum(-adf_h, 0.0)], axis=1)
    nh = np.stack([np.minimum(adf_h, 0.0), np.minimum(-adf_h, 0.0)], axis=1)
    fct_plus = pos_v + scatter_add(ph, mesh.edges, mesh.nod2D)
    fct_minus = neg_v + scatter_add(nh, mesh.edges, mesh.nod2D)
    fct_plus = _exch(fct_plus, "nod")
    fct_minus = _exch(fct_minus, "nod")

    a = mesh.areasvol
    safe_a = np.where(a > 0.0, a, 1.0)
    safe_hn = np.where(hnode_new > 0.0, hnode_new, 1.0)
    flux_pos = fct_plus * dt / safe_a / safe_hn + _FCT_FLUX_EPS
    flux_neg = fct_minus * dt / safe_a / safe_hn - _FCT_FLUX_EPS
    plus_fac = np.minimum(1.0, fct_ttf_max / flux_pos)
    minus_fac = np.minimum(1.0, fct_ttf_min / flux_neg)
    valid = nlm & (a > 0.0) & (hnode_new > 0.0)
    plus_fac = np.where(valid, plus_fac, 1.0)
    minus_fac = np.where(valid, minus_fac, 1.0)

    ae_surf = np.where(adf_v >= 0.0, plus_fac, minus_fac)
    ae_int = np.where(adf_v >= 0.0,
                      np.minimum(_shift_down(minus_fac), plus_fac),
                      np.minimum(_shift_down(plus_fac), minus_fac))
    ae_v = np.where(k == nzmin, ae_surf, ae_int)
    adf_v_lim = mask_below_bottom(ae_v * adf_v, mesh.node_iface_mask)
    n1, n2 = mesh.edges[:, 0], mesh.edges[:, 1]
    ae_h = np.where(adf_h >= 0.0,
                    np.minimum(plus_fac[n1], minus_fac[n2]),
                    np.minimum(minus_fac[n1], plus_fac[n2]))
    adf_h_lim = ae_h * adf_h
    if return_factors:
        return adf_h_lim, adf_v_lim, (plus_fac, minus_fac, fct_ttf_max, fct_ttf_min,
                                      fct_plus, fct_minus)
    return adf_h_lim, adf_v_lim


def flux2dtracer_fct(mesh, T, LO, adf_h, adf_v, hnode, hnode_new, *, dt=DT_DEFAULT):
    """fesom_jax/tracer_adv.py:469-484."""
    nlm = mesh.node_layer_mask
    a = mesh.areasvol
    safe_a = np.where(a > 0.0, a, 1.0)
    lo_trans = -T * hnode + LO * hnode_new
    vdiv = np.where(a > 0.0, (adf_v - _shift_up(adf_v)) * dt / safe_a, 0.0)
    dttf_v = np.where(nlm, lo_trans + vdiv, 0.0)
    hvals = np.stack([adf_h, -adf_h], axis=1)
    hsc = scatter_add(hvals, mesh.edges, mesh.nod2D)
    dttf_h = np.where(nlm, hsc * dt / safe_a, 0.0)
    return dttf_h + dttf_v


def adv_tra_vert_impl(mesh, w_i, T, hnode_new, *, dt=DT_DEFAULT):
    """Implicit vertical advection by ``w_i`` (fesom_jax/tracer_adv.py:487-529)."""
    nl = mesh.nl
    safe_av = np.where(mesh.areasvol > 0.0, mesh.areasvol, 1.0)
    v_top = dt * mesh.area / safe_av
    v_bot = dt * _shift_up(mesh.area) / safe_av
    wu = w_i
    wd = _shift_up(w_i)
    k = np.arange(nl)[None, :]
    nzmin = (mesh.ulevels_nod2D.astype(np.int64) - 1)[:, None]
    nzmax = (mesh.nlevels_nod2D.astype(np.int64) - 1)[:, None]
    valid = mesh.node_layer_mask
    surf = k == nzmin
    bot = k == (nzmax - 1)
    a_adv = np.where(surf, 0.0, np.minimum(0.0, wu)) * v_top
    b_top = np.where(surf, wu, np.maximum(0.0, wu)) * v_top
    b_bot = np.where(bot, 0.0, -np.minimum(0.0, wd)) * v_bot
    c_adv = np.where(bot, 0.0, -np.maximum(0.0, wd)) * v_bot
    b_adv = b_top + b_bot
    a = np.where(valid, a_adv, 0.0)
    c = np.where(valid, c_adv, 0.0)
    b = np.where(valid, hnode_new + b_adv, 1.0)
    rhs = np.where(valid, -(a * _shift_down(T) + b_adv * T + c * _shift_up(T)), 0.0)
    dT = tdma(a, b, c, rhs)
    return T + mask_below_bottom(dT, valid)


def advect_one_fct(mesh, uv, w_e, helem, hnode, hnode_new, T, T_old, *, dt=DT_DEFAULT,
                   exch=None, Z3d=None, zbar3=None, w_i=None, return_parts=False):
    """FCT driver, fesom_jax/tracer_adv.py:532-578. ``return_parts`` adds a dict of
    the intermediates (LO, tr_xy, adf_h, adf_v, limited fluxes, factors)."""
    _exch = exch if exch is not None else (lambda f, kind: f)
    T = np.asarray(T, np.float64)
    T_old = np.asarray(T_old, np.float64)
    ttfAB, T_old_new = tracer_ab(T, T_old)
    flux_h_lo = adv_flux_hor(mesh, uv, helem, T)
    flux_v_lo = adv_flux_ver(mesh, w_e, T)
    LO = compute_fct_lo(mesh, flux_h_lo, flux_v_lo, T, hnode, hnode_new, dt=dt)
    if w_i is not None:
        LO = adv_tra_vert_impl(mesh, w_i, LO, hnode_new, dt=dt)
    LO = _exch(LO, "nod")
    tr_xy = tracer_gradient_elements(mesh, T)
    tr_xy = _exch(tr_xy, "elem")
    eud = fill_up_dn_grad(mesh, tr_xy)
    adf_h = adv_flux_hor_ho(mesh, uv, helem, ttfAB, eud) - flux_h_lo
    if w_i is None:
        adf_v = adv_flux_ver_ho(mesh, w_e, ttfAB, Z3d=Z3d, zbar3=zbar3) - flux_v_lo
    else:
        w_full = w_e + w_i
        adf_v = (adv_flux_ver_ho(mesh, w_full, ttfAB, Z3d=Z3d, zbar3=zbar3)
                 - adv_flux_ver(mesh, w_full, T))
    adf_h_lim, adf_v_lim, fac = zalesak_limit(mesh, T, LO, adf_h, adf_v, hnode_new, dt=dt,
                                              exch=exch, return_factors=True)
    del_ttf = flux2dtracer_fct(mesh, T, LO, adf_h_lim, adf_v_lim, hnode, hnode_new, dt=dt)
    T_new = ale_reconstruct(mesh, T, del_ttf, hnode, hnode_new)
    if return_parts:
        parts = dict(LO=LO, tr_xy=tr_xy, eud=eud, adf_h=adf_h, adf_v=adf_v,
                     adf_h_lim=adf_h_lim, adf_v_lim=adf_v_lim, plus_fac=fac[0],
                     minus_fac=fac[1], fct_ttf_max=fac[2], fct_ttf_min=fac[3],
                     fct_plus=fac[4], fct_minus=fac[5], del_ttf=del_ttf,
                     flux_h_lo=flux_h_lo, flux_v_lo=flux_v_lo)
        return T_new, T_old_new, parts
    return T_new, T_old_new


# --------------------------------------------------------------------------
# input-side helpers (used to build test inputs)
# --------------------------------------------------------------------------
def compute_w(mesh, uv, helem):
    """Vertical velocity by continuity (fesom_jax/ale.py:309-365): per-level
    antisymmetric edge divergence, bottom-to-top cumulative sum, / area."""
    et = mesh.edge_tri
    el1, el2 = et[:, 0], et[:, 1]
    has1, has2 = el1 >= 0, el2 >= 0
    el1s = np.where(has1, el1, 0)
    el2s = np.where(has2, el2, 0)
    cross = mesh.edge_cross_dxdy
    U, V = uv[:, :, 0], uv[:, :, 1]
    lm = mesh.elem_layer_mask

    def cterm(els, has, dxcol, dycol, sign):
        dx = cross[:, dxcol:dxcol + 1]
        dy = cross[:, dycol:dycol + 1]
        m = lm[els] & has[:, None]
        return sign * np.where(m, (V[els] * dx - U[els] * dy) * helem[els], 0.0)

    c = cterm(el1s, has1, 0, 1, 1.0) + cterm(el2s, has2, 2, 3, -1.0)
    div = scatter_add(np.stack([c, -c], axis=1), mesh.edges, mesh.nod2D)
    w = np.cumsum(div[:, ::-1], axis=1)[:, ::-1]
    w = w / np.where(mesh.area > 0.0, mesh.area, 1.0)
    return mask_below_bottom(w, mesh.node_iface_mask)
