"""CPU oracle for the adjoint (test infrastructure only): a hand-written reverse-mode
restatement of one FCT step, fesom_jax/tracer_adv.py:532-578 (with or without the
implicit split ``w_i`` and the moving-level stencil ``Z3d/zbar3``), giving the
cotangents of ``T, T_old, uv, w_e`` (and ``w_i``) for a cotangent of ``T_new``.
``Z3d``, ``zbar3``, ``helem``, ``hnode`` and ``hnode_new`` are treated as constants.

It follows the differentiation conventions of ``jax.grad`` applied to the reference
(fesom_jax/docs/LIMITER_GRADIENTS.md:45-75: plain subgradients, no custom VJP):

* ``max(x, y)`` / ``min(x, y)``: the cotangent goes to the larger / smaller argument,
  half to each on an exact tie;
* a max/min reduction over an axis: shared equally by the tied arguments;
* ``segment_max`` / ``segment_min``: shared equally by the updates equal to the
  result (the segment's -inf/+inf identity never ties a finite value);
* ``abs``: derivative +1 at 0; ``where``: the chosen branch only;
* ``x / y``: ``g / y`` to x and ``(-g * x) * (1 / (y * y))`` to y.

Only ``tests/`` and ``bench.py``'s baseline leg use this module. The tie conventions
are restated from JAX's published rules and cannot be run against JAX here (it is not
installed): for exact ties they are "parity unpinned"; at generic points (no ties)
the cotangents are pinned by finite differences and by the transpose identity against
the oracle's forward (tests/test_oracle_vjp.py).
"""

from __future__ import annotations

import numpy as np

from . import fct_numpy as F

R_EARTH = F.R_EARTH


def _sd_T(g):
    """Transpose of _shift_down (out[k] = x[k-1], out[0] = x[0])."""
    out = np.zeros_like(g)
    out[..., :-1] += g[..., 1:]
    out[..., 0] += g[..., 0]
    return out


def _su_T(g):
    """Transpose of _shift_up (out[k] = x[k+1], 0 at the end)."""
    out = np.zeros_like(g)
    out[..., 1:] += g[..., :-1]
    return out


def _max_vjp(g, x, y):
    """Cotangents of max(x, y)."""
    gx = np.where(x > y, g, np.where(x == y, 0.5 * g, 0.0))
    gy = np.where(y > x, g, np.where(x == y, 0.5 * g, 0.0))
    return gx, gy


def _min_vjp(g, x, y):
    gx = np.where(x < y, g, np.where(x == y, 0.5 * g, 0.0))
    gy = np.where(y < x, g, np.where(x == y, 0.5 * g, 0.0))
    return gx, gy


def _abs_vjp(g, x):
    return np.where(x >= 0.0, g, -g)


def _div_vjp(g, x, y):
    return g / y, (-g * x) * (1.0 / (y * y))


def _scatter_T(g_node, seg):
    """Transpose of the masked segment sum: the per-entry cotangent is a gather."""
    seg = np.asarray(seg)
    safe = np.where(seg >= 0, seg, 0)
    out = g_node[safe]
    return np.where((seg >= 0).reshape(seg.shape + (1,) * (out.ndim - seg.ndim)), out, 0.0)


def _vimpl_rows(mesh, w_i, hnode_new, dt):
    """Row coefficients of adv_tra_vert_impl (fesom_jax/tracer_adv.py:487-529) and the
    pieces their cotangents need."""
    m = mesh
    L = m.nl
    safe_av = np.where(m.areasvol > 0.0, m.areasvol, 1.0)
    v_top = dt * m.area / safe_av
    v_bot = dt * F._shift_up(m.area) / safe_av
    wu, wd = w_i, F._shift_up(w_i)
    k = np.arange(L)[None, :]
    nzmin = (m.ulevels_nod2D.astype(np.int64) - 1)[:, None]
    nzmax = (m.nlevels_nod2D.astype(np.int64) - 1)[:, None]
    valid = m.node_layer_mask
    surf = k == nzmin
    bot = k == (nzmax - 1)
    b_adv = (np.where(surf, wu, np.maximum(0.0, wu)) * v_top
             + np.where(bot, 0.0, -np.minimum(0.0, wd)) * v_bot)
    a = np.where(valid, np.where(surf, 0.0, np.minimum(0.0, wu)) * v_top, 0.0)
    c = np.where(valid, np.where(bot, 0.0, -np.maximum(0.0, wd)) * v_bot, 0.0)
    b = np.where(valid, hnode_new + b_adv, 1.0)
    return dict(a=a, b=b, c=c, b_adv=b_adv, v_top=v_top, v_bot=v_bot, wu=wu, wd=wd,
                surf=surf, bot=bot, valid=valid)


def _vimpl_vjp(mesh, w_i, hnode_new, dt, X, g_out):
    """Cotangents ``(g_X, g_w_i)`` of ``adv_tra_vert_impl(mesh, w_i, X, hnode_new)``
    (fesom_jax/tracer_adv.py:487-529) for ``g_out``: reverse mode through the Thomas
    recurrence of ops.tdma (fesom_jax/ops.py:121-168) step by step, as jax.grad
    differentiates its two scans."""
    r = _vimpl_rows(mesh, w_i, hnode_new, dt)
    a, b, c, b_adv, valid = r["a"], r["b"], r["c"], r["b_adv"], r["valid"]
    rhs = np.where(valid, -(a * F._shift_down(X) + b_adv * X + c * F._shift_up(X)), 0.0)
    N, L = X.shape
    cp = np.zeros((N, L))
    dp = np.zeros((N, L))
    mm = np.zeros((N, L))
    cprev = np.zeros(N)
    dprev = np.zeros(N)
    for i in range(L):
        mm[:, i] = b[:, i] - a[:, i] * cprev
        cp[:, i] = c[:, i] / mm[:, i]
        dp[:, i] = (rhs[:, i] - a[:, i] * dprev) / mm[:, i]
        cprev, dprev = cp[:, i], dp[:, i]
    x = np.zeros((N, L))
    xn = np.zeros(N)
    for i in range(L - 1, -1, -1):
        x[:, i] = dp[:, i] - cp[:, i] * xn
        xn = x[:, i]
    # out = X + mask(x): the back substitution reversed (x[i] feeds x[i-1])
    g_x = np.where(valid, g_out, 0.0)
    g_cp = np.zeros((N, L))
    g_dp = np.zeros((N, L))
    carry = np.zeros(N)
    for i in range(L):
        G = g_x[:, i] + carry
        g_dp[:, i] = G
        g_cp[:, i] = -G * (x[:, i + 1] if i + 1 < L else 0.0)
        carry = -G * cp[:, i]
    # the forward sweep reversed
    g_a = np.zeros((N, L))
    g_b = np.zeros((N, L))
    g_c = np.zeros((N, L))
    g_rhs = np.zeros((N, L))
    gcc = np.zeros(N)
    gdc = np.zeros(N)
    for i in range(L - 1, -1, -1):
        GC = g_cp[:, i] + gcc
        GD = g_dp[:, i] + gdc
        cprev = cp[:, i - 1] if i > 0 else np.zeros(N)
        dprev = dp[:, i - 1] if i > 0 else np.zeros(N)
        num = rhs[:, i] - a[:, i] * dprev
        g_num = GD / mm[:, i]
        g_m = (-GD * num) * (1.0 / (mm[:, i] * mm[:, i]))
        g_rhs[:, i] = g_num
        g_a[:, i] += -g_num * dprev
        gdc = -g_num * a[:, i]
        g_c[:, i] += GC / mm[:, i]
        g_m = g_m + (-GC * c[:, i]) * (1.0 / (mm[:, i] * mm[:, i]))
        g_b[:, i] += g_m
        g_a[:, i] += -g_m * cprev
        gcc = -g_m * a[:, i]
    # rhs = where(valid, -(a X[k-1] + b_adv X + c X[k+1]), 0)
    gs = -np.where(valid, g_rhs, 0.0)
    g_X = g_out.copy()
    g_a += gs * F._shift_down(X)
    g_badv = gs * X
    g_c += gs * F._shift_up(X)
    g_X += _sd_T(gs * a) + gs * b_adv + _su_T(gs * c)
    g_badv += np.where(valid, g_b, 0.0)
    g_aadv = np.where(valid, g_a, 0.0) * r["v_top"]
    g_cadv = np.where(valid, g_c, 0.0) * r["v_bot"]
    surf, bot, wu, wd = r["surf"], r["bot"], r["wu"], r["wd"]
    g_wu = np.where(surf, 0.0, _min_vjp(g_aadv, wu, np.zeros_like(wu))[0])
    gt = g_badv * r["v_top"]
    g_wu += np.where(surf, gt, _max_vjp(gt, wu, np.zeros_like(wu))[0])
    gb = g_badv * r["v_bot"]
    g_wd = np.where(bot, 0.0, -_min_vjp(gb, wd, np.zeros_like(wd))[0])
    g_wd += np.where(bot, 0.0, -_max_vjp(g_cadv, wd, np.zeros_like(wd))[0])
    return g_X, g_wu + _su_T(g_wd)


def forward_tape(mesh, uv, w_e, helem, hnode, hnode_new, T, T_old, dt, Z3d=None, zbar3=None,
                 w_i=None):
    """The forward step with every intermediate the backward needs."""
    m = mesh
    t = dict(T=T, T_old=T_old)
    A, _ = F.tracer_ab(T, T_old)
    t["A"] = A
    vf = F.edge_vflux(m, uv, helem)
    t["vf"] = vf
    n1, n2 = m.edges[:, 0], m.edges[:, 1]
    av = np.abs(vf)
    T1, T2 = T[n1], T[n2]
    fh = -0.5 * (T1 * (vf + av) + T2 * (vf - av))
    fv = F.adv_flux_ver(m, w_e, T)
    t.update(fh=fh, fv=fv)
    LO = F.compute_fct_lo(m, fh, fv, T, hnode, hnode_new, dt=dt)
    if w_i is not None:
        t["LOx"] = LO
        LO = F.adv_tra_vert_impl(m, w_i, LO, hnode_new, dt=dt)
    t["LO"] = LO
    tr = F.tracer_gradient_elements(m, T)
    gnode = F.node_avg_grad(m, tr)
    eud = F.fill_up_dn_grad(m, tr)
    t.update(tr=tr, gnode=gnode, eud=eud)
    ho = F.adv_flux_hor_ho(m, uv, helem, A, eud)
    adf_h = ho - fh
    if w_i is None:
        adf_v = F.adv_flux_ver_ho(m, w_e, A, Z3d=Z3d, zbar3=zbar3) - fv
    else:
        w_full = w_e + w_i
        adf_v = (F.adv_flux_ver_ho(m, w_full, A, Z3d=Z3d, zbar3=zbar3)
                 - F.adv_flux_ver(m, w_full, T))
    t.update(adf_h=adf_h, adf_v=adf_v)
    adf_h_lim, adf_v_lim, fac = F.zalesak_limit(m, T, LO, adf_h, adf_v, hnode_new, dt=dt,
                                                return_factors=True)
    t.update(adf_h_lim=adf_h_lim, adf_v_lim=adf_v_lim, pf=fac[0], mf=fac[1],
             fct_ttf_max=fac[2], fct_ttf_min=fac[3], fct_plus=fac[4], fct_minus=fac[5])
    del_ttf = F.flux2dtracer_fct(m, T, LO, adf_h_lim, adf_v_lim, hnode, hnode_new, dt=dt)
    T_new = F.ale_reconstruct(m, T, del_ttf, hnode, hnode_new)
    t["T_new"] = T_new
    return t


def advect_one_fct_vjp(mesh, uv, w_e, helem, hnode, hnode_new, T, T_old, g_Tnew, *,
                       dt=F.DT_DEFAULT, Z3d=None, zbar3=None, w_i=None):
    """Cotangents ``(g_T, g_T_old, g_uv, g_w_e)`` of one FCT step for ``g_Tnew``, and
    ``g_w_i`` as a fifth entry when ``w_i`` is given."""
    m = mesh
    L = m.nl
    tp = forward_tape(m, uv, w_e, helem, hnode, hnode_new, T, T_old, dt, Z3d=Z3d,
                      zbar3=zbar3, w_i=w_i)
    nlm = m.node_layer_mask
    nim = m.node_iface_mask
    a = m.areasvol
    safe_a = np.where(a > 0.0, a, 1.0)
    safe_hn = np.where(hnode_new > 0.0, hnode_new, 1.0)
    n1, n2 = m.edges[:, 0], m.edges[:, 1]
    k = np.arange(L)[None, :]
    nzmin = (m.ulevels_nod2D.astype(np.int64) - 1)[:, None]
    nzmax = (m.nlevels_nod2D.astype(np.int64) - 1)[:, None]
    N, E, D = m.nod2D, m.elem2D, m.edge2D

    g_T = np.zeros((N, L))
    g_A = np.zeros((N, L))
    g_LO = np.zeros((N, L))
    g_vf = np.zeros((D, L))
    g_w = np.zeros((N, L))

    # ---- ale_reconstruct (tracer_adv.py:142-151)
    g_T += g_Tnew
    upd = nlm & (hnode_new > 0.0)
    g_del = np.where(upd, g_Tnew / safe_hn, 0.0)
    g_T += g_del * (hnode - hnode_new)
    # ---- flux2dtracer_fct (tracer_adv.py:469-484)
    g_in = np.where(nlm, g_del, 0.0)
    g_T += -(g_in * hnode)
    g_LO += g_in * hnode_new
    g_x = np.where(a > 0.0, g_in, 0.0) / safe_a * dt
    g_vlim = g_x + (-_su_T(g_x))
    g_hsc = g_in / safe_a * dt
    g_hlim = _scatter_T(g_hsc, m.edges)                 # [D, 2, L]
    g_hlim = g_hlim[:, 0] - g_hlim[:, 1]
    # ---- zalesak_limit b3 (tracer_adv.py:453-466)
    adf_h, adf_v = tp["adf_h"], tp["adf_v"]
    pf, mf = tp["pf"], tp["mf"]
    g_pf = np.zeros((N, L))
    g_mf = np.zeros((N, L))
    g_adfh = g_hlim * np.minimum(np.where(adf_h >= 0, pf[n1], mf[n1]),
                                 np.where(adf_h >= 0, mf[n2], pf[n2]))
    g_ae = g_hlim * adf_h
    pos = adf_h >= 0.0
    ga1, ga2 = _min_vjp(g_ae, np.where(pos, pf[n1], mf[n1]), np.where(pos, mf[n2], pf[n2]))
    np.add.at(g_pf, n1, np.where(pos, ga1, 0.0))
    np.add.at(g_mf, n1, np.where(pos, 0.0, ga1))
    np.add.at(g_mf, n2, np.where(pos, ga2, 0.0))
    np.add.at(g_pf, n2, np.where(pos, 0.0, ga2))
    g_vl = np.where(nim, g_vlim, 0.0)
    sdm, sdp = F._shift_down(mf), F._shift_down(pf)
    ae_surf = np.where(adf_v >= 0.0, pf, mf)
    ae_int = np.where(adf_v >= 0.0, np.minimum(sdm, pf), np.minimum(sdp, mf))
    ae_v = np.where(k == nzmin, ae_surf, ae_int)
    g_adfv = g_vl * ae_v
    g_aev = g_vl * adf_v
    surf = k == nzmin
    pv = adf_v >= 0.0
    g_s = np.where(surf, g_aev, 0.0)
    g_pf += np.where(pv, g_s, 0.0)
    g_mf += np.where(pv, 0.0, g_s)
    g_i = np.where(surf, 0.0, g_aev)
    x1, y1 = _min_vjp(np.where(pv, g_i, 0.0), sdm, pf)        # adf_v >= 0 branch
    x2, y2 = _min_vjp(np.where(pv, 0.0, g_i), sdp, mf)
    g_mf += _sd_T(x1)
    g_pf += y1
    g_pf += _sd_T(x2)
    g_mf += y2
    # ---- zalesak_limit b2 (tracer_adv.py:441-451)
    valid = nlm & (a > 0.0) & (hnode_new > 0.0)
    fmax_, fmin_ = tp["fct_ttf_max"], tp["fct_ttf_min"]
    fp, fm = tp["fct_plus"], tp["fct_minus"]
    flux_pos = fp * dt / safe_a / safe_hn + F._FCT_FLUX_EPS
    flux_neg = fm * dt / safe_a / safe_hn - F._FCT_FLUX_EPS
    rp = fmax_ / flux_pos
    rm = fmin_ / flux_neg
    _, g_rp = _min_vjp(np.where(valid, g_pf, 0.0), np.ones_like(rp), rp)
    _, g_rm = _min_vjp(np.where(valid, g_mf, 0.0), np.ones_like(rm), rm)
    g_fmax, g_flp = _div_vjp(g_rp, fmax_, flux_pos)
    g_fmin, g_fln = _div_vjp(g_rm, fmin_, flux_neg)
    g_fp = g_flp / safe_hn / safe_a * dt
    g_fm = g_fln / safe_hn / safe_a * dt
    # ---- zalesak_limit b1 (tracer_adv.py:426-435)
    gp_e = _scatter_T(g_fp, m.edges)
    gm_e = _scatter_T(g_fm, m.edges)
    ph0, ph1 = _max_vjp(gp_e[:, 0], adf_h, np.zeros_like(adf_h))[0], \
        _max_vjp(gp_e[:, 1], -adf_h, np.zeros_like(adf_h))[0]
    nh0, nh1 = _min_vjp(gm_e[:, 0], adf_h, np.zeros_like(adf_h))[0], \
        _min_vjp(gm_e[:, 1], -adf_h, np.zeros_like(adf_h))[0]
    g_adfh += ph0 - ph1 + nh0 - nh1
    fvb = F._shift_up(adf_v)
    gpv = np.where(nlm, g_fp, 0.0)
    gnv = np.where(nlm, g_fm, 0.0)
    g_adfv += _max_vjp(gpv, adf_v, np.zeros_like(adf_v))[0]
    g_adfv += _min_vjp(gnv, adf_v, np.zeros_like(adf_v))[0]
    g_fvb = -_max_vjp(gpv, -fvb, np.zeros_like(fvb))[0] - _min_vjp(gnv, -fvb, np.zeros_like(fvb))[0]
    g_adfv += _su_T(g_fvb)
    # ---- zalesak_limit a1-a4 (tracer_adv.py:389-424)
    LO = tp["LO"]
    g_fmax = np.where(nlm, g_fmax, 0.0)
    g_fmin = np.where(nlm, g_fmin, 0.0)
    g_LO -= g_fmax + g_fmin
    use_own = (k == nzmin) | (k == nzmax - 1)
    ttf_max = np.maximum(LO, T)
    ttf_min = np.minimum(LO, T)
    em = m.elem_layer_mask
    en = m.elem_nodes
    vmax3 = ttf_max[en]                                 # [E, 3, L]
    vmin3 = ttf_min[en]
    aux_max = np.where(em, vmax3.max(axis=1), -F._FCT_BIG)
    aux_min = np.where(em, vmin3.min(axis=1), F._FCT_BIG)
    seg = en.reshape(-1)
    cmax = np.broadcast_to(aux_max[:, None], (E, 3, L)).reshape(3 * E, L)
    cmin = np.broadcast_to(aux_min[:, None], (E, 3, L)).reshape(3 * E, L)
    tvx = F.segment_max(cmax, seg, N)
    tvn = F.segment_min(cmin, seg, N)

    def cluster_T(g, tv, op_vjp):
        g_tv = np.where(use_own, g, 0.0)
        g3 = np.where(use_own, 0.0, g)
        a_, b_, c_ = F._shift_down(tv), tv, F._shift_up(tv)
        m1 = np.maximum(a_, b_) if op_vjp is _max_vjp else np.minimum(a_, b_)
        gm1, gc = op_vjp(g3, m1, c_)
        ga_, gb_ = op_vjp(gm1, a_, b_)
        return g_tv + _sd_T(ga_) + gb_ + _su_T(gc)

    g_tvx = cluster_T(g_fmax, tvx, _max_vjp)
    g_tvn = cluster_T(g_fmin, tvn, _min_vjp)

    def seg_T(g_tv, c, tv):
        """segment extremum: share among tied updates."""
        hit = c == tv[seg]
        cnt = np.zeros((N, L))
        np.add.at(cnt, seg, hit.astype(float))
        share = np.where(hit, g_tv[seg] / np.where(cnt[seg] > 0, cnt[seg], 1.0), 0.0)
        return share.reshape(E, 3, L).sum(axis=1)

    g_auxx = seg_T(g_tvx, cmax, tvx)
    g_auxn = seg_T(g_tvn, cmin, tvn)

    def reduce_T(g_aux, v3, aux):
        g = np.where(em, g_aux, 0.0)
        hit = v3 == aux[:, None, :]
        cnt = hit.sum(axis=1, keepdims=True)
        return np.where(hit, g[:, None, :] / np.where(cnt > 0, cnt, 1), 0.0)

    gv_max = reduce_T(g_auxx, vmax3, aux_max)            # [E, 3, L]
    gv_min = reduce_T(g_auxn, vmin3, aux_min)
    g_tmax = np.zeros((N, L))
    g_tmin = np.zeros((N, L))
    np.add.at(g_tmax, seg, gv_max.reshape(3 * E, L))
    np.add.at(g_tmin, seg, gv_min.reshape(3 * E, L))
    gl, gt = _max_vjp(g_tmax, LO, T)
    g_LO += gl
    g_T += gt
    gl, gt = _min_vjp(g_tmin, LO, T)
    g_LO += gl
    g_T += gt
    # ---- adf_v = adv_flux_ver_ho(A) - adv_flux_ver(T) (tracer_adv.py:241-295, 108-120)
    g_hov = g_adfv
    g_fv = -g_adfv      # the LO-form flux subtracted in adf_v (with w_e + w_i if split)
    W = w_e if w_i is None else w_e + w_i
    area, A = m.area, tp["A"]
    g_wf = np.zeros((N, L))     # cotangent of the w in adf_v
    g_hov = np.where(nim, g_hov, 0.0)
    is_surf = k == nzmin
    is_bot = k == nzmax
    is_cent = (k == nzmin + 1) | (k == nzmax - 1)
    is_int = (k >= nzmin + 2) & (k <= nzmax - 2)
    gs_ = np.where(is_surf, g_hov, 0.0)
    gc_ = np.where(~is_surf & ~is_bot & is_cent, g_hov, 0.0)
    gq_ = np.where(~is_surf & ~is_bot & ~is_cent & is_int, g_hov, 0.0)
    # surf = ((-A) * W) * area
    g_A += -(gs_ * area * W)
    g_wf += gs_ * area * (-A)
    # cent = ((-0.5 * (Am1 + A)) * W) * area
    Am1 = F._shift_down(A)
    Am2 = F._shift_down(Am1)
    Ap1 = F._shift_up(A)
    gcm = gc_ * area
    g_sum = -0.5 * (gcm * W)
    g_wf += gcm * (-0.5 * (Am1 + A))
    g_A += g_sum
    g_Am1 = g_sum.copy()
    # quad
    if Z3d is None:
        Z_nz, Z_m1, Z_m2, Z_p1 = F._z_stencil(np.asarray(m.Z), L)
        zb = np.asarray(m.zbar)
    else:
        Z_nz = Z3d
        Z_m1 = F._shift_down(Z3d)
        Z_m2 = F._shift_down(Z_m1)
        Z_p1 = np.concatenate([Z3d[:, 1:], Z3d[:, -1:]], axis=1)
        zb = zbar3
    sf = lambda x: np.where(x == 0.0, 1.0, x)
    d_c, d_u, d_d = sf(Z_m1 - Z_nz), sf(Z_nz - Z_p1), sf(Z_m2 - Z_m1)
    with np.errstate(divide="ignore", invalid="ignore"):
        qc = (Am1 - A) / d_c
        qu = (A - Ap1) / d_u
        qd = (Am2 - Am1) / d_d
        Tmean1 = A + (2.0 * qc + qu) * (zb - Z_nz) / 3.0
        Tmean2 = Am1 + (2.0 * qc + qd) * (zb - Z_m1) / 3.0
    gqm = gq_ * area
    g_wf += np.where(is_int, gqm * (-0.5 * (Tmean1 + Tmean2)), 0.0)
    g_tm = -0.5 * (gqm * W)
    g_tm = np.where(is_int, g_tm, 0.0)
    g_A += g_tm
    g_Am1 += g_tm
    g_r1 = g_tm / 3.0 * (zb - Z_nz)
    g_r2 = g_tm / 3.0 * (zb - Z_m1)
    g_qc = 2.0 * g_r1 + 2.0 * g_r2
    g_qu = g_r1
    g_qd = g_r2
    g_Am1 += g_qc / d_c
    g_A -= g_qc / d_c
    g_A += g_qu / d_u
    g_Ap1 = -(g_qu / d_u)
    g_Am2 = g_qd / d_d
    g_Am1 -= g_qd / d_d
    g_Am1 += _sd_T(g_Am2)
    g_A += _sd_T(g_Am1) + _su_T(g_Ap1)
    # ---- adf_h = adv_flux_hor_ho(A, eud) - fh (tracer_adv.py:196-225)
    vf = tp["vf"]
    av = np.abs(vf)
    eud = tp["eud"]
    et = m.edge_tri
    el1, el2 = et[:, 0], et[:, 1]
    has2 = el2 >= 0
    el1s = np.where(el1 >= 0, el1, 0)
    el2s = np.where(has2, el2, 0)
    aa = R_EARTH * m.elem_cos[el1s]
    aa = np.where(has2, 0.5 * (aa + R_EARTH * m.elem_cos[el2s]), aa)[:, None]
    edx = m.edge_dxdy[:, 0:1]
    edy = m.edge_dxdy[:, 1:2]
    A1, A2 = A[n1], A[n2]
    diff = A2 - A1
    g0, g1, g2, g3 = eud[:, :, 0], eud[:, :, 1], eud[:, :, 2], eud[:, :, 3]
    Tm1 = A1 + (2.0 * diff + edx * aa * g0 + edy * R_EARTH * g2) / 6.0
    Tm2 = A2 - (2.0 * diff + edx * aa * g1 + edy * R_EARTH * g3) / 6.0
    g_ho = g_adfh
    g_fh = -g_adfh
    gs2 = -0.5 * g_ho
    g_Tm1 = gs2 * (vf + av)
    g_Tm2 = gs2 * (vf - av)
    g_vp = gs2 * Tm1
    g_vm = gs2 * Tm2
    g_vf += g_vp + g_vm + _abs_vjp(g_vp - g_vm, vf)
    gi1 = g_Tm1 / 6.0
    gi2 = -(g_Tm2 / 6.0)
    g_A1 = g_Tm1 - 2.0 * (gi1 + gi2)
    g_A2 = g_Tm2 + 2.0 * (gi1 + gi2)
    g_eud = np.stack([gi1 * (edx * aa), gi2 * (edx * aa), gi1 * (edy * R_EARTH),
                      gi2 * (edy * R_EARTH)], axis=-1)
    np.add.at(g_A, n1, g_A1)
    np.add.at(g_A, n2, g_A2)
    # ---- fill_up_dn_grad / _node_avg_grad / tracer_gradient_elements (:316-368)
    up, dn = m.edge_up_dn_tri[:, 0], m.edge_up_dn_tri[:, 1]
    interior = (up >= 0) & (dn >= 0)
    nzmin_e = np.maximum(m.ulevels_nod2D_max[n1], m.ulevels_nod2D_max[n2]).astype(np.int64)[:, None] - 1
    nzmax_e = np.minimum(m.nlevels_nod2D_min[n1], m.nlevels_nod2D_min[n2]).astype(np.int64)[:, None] - 1
    use_tri = interior[:, None] & (k >= nzmin_e) & (k < nzmax_e)
    g_tr = np.zeros((E, L, 2))
    ups = np.where(up >= 0, up, 0)
    dns = np.where(dn >= 0, dn, 0)
    np.add.at(g_tr, ups, np.stack([np.where(use_tri, g_eud[..., 0], 0.0),
                                   np.where(use_tri, g_eud[..., 2], 0.0)], axis=-1))
    np.add.at(g_tr, dns, np.stack([np.where(use_tri, g_eud[..., 1], 0.0),
                                   np.where(use_tri, g_eud[..., 3], 0.0)], axis=-1))
    g_gn = np.zeros((N, L, 2))
    np.add.at(g_gn, n1, np.stack([np.where(use_tri, 0.0, g_eud[..., 0]),
                                  np.where(use_tri, 0.0, g_eud[..., 2])], axis=-1))
    np.add.at(g_gn, n2, np.stack([np.where(use_tri, 0.0, g_eud[..., 1]),
                                  np.where(use_tri, 0.0, g_eud[..., 3])], axis=-1))
    aw = np.where(em, m.elem_area[:, None], 0.0)
    den = np.zeros((N, L))
    np.add.at(den, seg, np.broadcast_to(aw[:, None], (E, 3, L)).reshape(3 * E, L))
    g_num = g_gn / np.where(den > 0.0, den, 1.0)[:, :, None]
    g_tr += g_num[en].sum(axis=1) * aw[:, :, None]
    g_tr = np.where(em[:, :, None], g_tr, 0.0)
    gs = m.gradient_sca
    g_tv3 = (g_tr[:, None, :, 0] * gs[:, 0:3, None] + g_tr[:, None, :, 1] * gs[:, 3:6, None])
    np.add.at(g_T, seg, g_tv3.reshape(3 * E, L))
    # ---- adv_tra_vert_impl applied to LO (tracer_adv.py:555-556, 487-529)
    g_wi = None
    if w_i is not None:
        g_LO, g_wi = _vimpl_vjp(m, w_i, hnode_new, dt, tp["LOx"], g_LO)
    # ---- LO (tracer_adv.py:298-313) and the low-order fluxes
    valid_lo = nlm & (hnode_new > 0.0) & (a > 0.0)
    g_num_lo = np.where(valid_lo, g_LO, 0.0) / safe_hn
    g_T += g_num_lo * hnode
    g_div = g_num_lo / safe_a * dt
    g_fh_e = _scatter_T(g_div, m.edges)
    g_fh += g_fh_e[:, 0] - g_fh_e[:, 1]
    g_fvlo = g_div + (-_su_T(g_div))
    if w_i is None:
        g_fv = g_fv + g_fvlo
    # fh = -0.5 * (T1 (vf + av) + T2 (vf - av))
    gs1 = -0.5 * g_fh
    g_T1 = gs1 * (vf + av)
    g_T2 = gs1 * (vf - av)
    g_vp = gs1 * T[n1]
    g_vm = gs1 * T[n2]
    g_vf += g_vp + g_vm + _abs_vjp(g_vp - g_vm, vf)
    np.add.at(g_T, n1, g_T1)
    np.add.at(g_T, n2, g_T2)
    # fv = mask(((-0.5 * (T (w + aw) + Tab (w - aw))) * area)), with w_e in LO and the
    # adf_v w (w_e + w_i when split) in adf_v
    def fv_T(g, Wv):
        g = np.where(nlm, g, 0.0) * area
        gs3 = -0.5 * g
        aw_ = np.abs(Wv)
        gt = gs3 * (Wv + aw_) + _sd_T(gs3 * (Wv - aw_))
        g_wp = gs3 * T
        g_wm = gs3 * F._shift_down(T)
        return gt, g_wp + g_wm + _abs_vjp(g_wp - g_wm, Wv)

    if w_i is None:
        gt, gw_ = fv_T(g_fv, w_e)
        g_T += gt
        g_w += gw_ + g_wf
    else:
        gt, gw_ = fv_T(g_fvlo, w_e)
        gt2, gwf_ = fv_T(g_fv, W)
        g_T += gt + gt2
        g_wf += gwf_
        g_w += gw_ + g_wf
        g_wi = g_wi + g_wf
    # ---- tracer_ab (tracer_adv.py:59-66)
    g_T += F._AB_NEW * g_A
    g_Told = F._AB_OLD * g_A
    # ---- edge_vflux (tracer_adv.py:69-93): cotangent of uv
    g_uv = np.zeros_like(uv)
    cross = m.edge_cross_dxdy
    lm = m.elem_layer_mask
    has1 = el1 >= 0
    m1 = lm[el1s] & has1[:, None]
    m2 = lm[el2s] & has2[:, None]
    gh1 = np.where(m1, g_vf, 0.0) * helem[el1s]
    gh2 = np.where(m2, g_vf, 0.0) * helem[el2s]
    np.add.at(g_uv[:, :, 0], el1s, gh1 * cross[:, 1:2])
    np.add.at(g_uv[:, :, 1], el1s, -(gh1 * cross[:, 0:1]))
    np.add.at(g_uv[:, :, 1], el2s, gh2 * cross[:, 2:3])
    np.add.at(g_uv[:, :, 0], el2s, -(gh2 * cross[:, 3:4]))
    if w_i is not None:
        return g_T, g_Told, g_uv, g_w, g_wi
    return g_T, g_Told, g_uv, g_w


def compute_w_vjp(mesh, helem, g_w):
    """Cotangent of uv for g_w through fesom_jax/ale.py:309-365 (linear map)."""
    m = mesh
    g = np.where(m.node_iface_mask, g_w, 0.0) / np.where(m.area > 0.0, m.area, 1.0)
    # reverse cumulative sum transposes to a forward cumulative sum
    g_div = np.cumsum(g, axis=1)
    ge = _scatter_T(g_div, m.edges)
    g_c = ge[:, 0] - ge[:, 1]
    et = m.edge_tri
    el1, el2 = et[:, 0], et[:, 1]
    has1, has2 = el1 >= 0, el2 >= 0
    el1s = np.where(has1, el1, 0)
    el2s = np.where(has2, el2, 0)
    cross = m.edge_cross_dxdy
    lm = m.elem_layer_mask
    g_uv = np.zeros((m.elem2D, m.nl, 2))
    g1 = np.where(lm[el1s] & has1[:, None], g_c, 0.0) * helem[el1s]
    g2 = np.where(lm[el2s] & has2[:, None], -g_c, 0.0) * helem[el2s]
    np.add.at(g_uv[:, :, 1], el1s, g1 * cross[:, 0:1])
    np.add.at(g_uv[:, :, 0], el1s, -(g1 * cross[:, 1:2]))
    np.add.at(g_uv[:, :, 1], el2s, g2 * cross[:, 2:3])
    np.add.at(g_uv[:, :, 0], el2s, -(g2 * cross[:, 3:4]))
    return g_uv
