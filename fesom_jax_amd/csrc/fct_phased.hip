// Phased FP64 FCT kernels for gfx950: one kernel per stage of the reference's
// advect_one_fct (fesom_jax/tracer_adv.py:532-578), each a 1-D grid over the
// flattened (entity, level) lanes of its output. Every edge->node and element->node
// sum is a gather over a per-node CSR list kept in the reference's scatter order
// (flattened (edge, side) / (element, vertex) index), accumulated from +0.0 in
// that order, so the result does not depend on scheduling and is reproducible bit
// for bit. The file is compiled with -ffp-contract=off: every expression below keeps
// the operation order of the reference's jnp expression it cites.
//
// This path also serves the caller-visible phases (the reference's exch hook).

#include "fct_common.h"

namespace fct {

// Low-order vertical flux at interface k (adv_flux_ver, tracer_adv.py:108-120):
// -0.5*(T[k](w+|w|) + T[k-1](w-|w|))*area with T[-1] := T[0]; 0 off the layer mask.
__device__ __forceinline__ double flux_v_lo(const DevMesh& m, const double* __restrict__ T,
                                            double wk, int n, int k, int2 nr) {
    if (!in_layer(nr, k)) return 0.0;
    const size_t i = (size_t)n * m.L + k;
    const double aw = fabs(wk);
    const double Tb = T[i];
    const double Ta = T[k > 0 ? i - 1 : i];
    return (-0.5 * (Tb * (wk + aw) + Ta * (wk - aw))) * m.area[i];
}

__device__ __forceinline__ double w_at(const double* __restrict__ w_e,
                                       const double* __restrict__ w_i, size_t i) {
    return w_i ? (w_e[i] + w_i[i]) : w_e[i];
}

// edge_vflux, tracer_adv.py:69-93.
__global__ void __launch_bounds__(256)
k_vflux(DevMesh m, const double* __restrict__ uv, const double* __restrict__ helem,
        double* __restrict__ vflux) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.D * m.L) return;
    const int e = (int)(t / m.L), k = (int)(t - (int64_t)e * m.L);
    const int2 et = m.edge_tri[e];
    const double4 c = m.cross[e];
    double t1 = 0.0, t2 = 0.0;
    if (et.x >= 0 && in_layer(m.elev[et.x], k)) {
        const size_t i = (size_t)et.x * m.L + k;
        const double U = uv[2 * i], V = uv[2 * i + 1];
        t1 = (U * c.y - V * c.x) * helem[i];
    }
    if (et.y >= 0 && in_layer(m.elev[et.y], k)) {
        const size_t i = (size_t)et.y * m.L + k;
        const double U = uv[2 * i], V = uv[2 * i + 1];
        t2 = (V * c.z - U * c.w) * helem[i];
    }
    vflux[t] = t1 + t2;
}

// compute_fct_lo with the LO fluxes of adv_flux_hor / adv_flux_ver on values T
// (tracer_adv.py:96-120, 298-313).
__global__ void __launch_bounds__(256)
k_lo(DevMesh m, const double* __restrict__ vflux, const double* __restrict__ T,
     const double* __restrict__ w_e, const double* __restrict__ hnode,
     const double* __restrict__ hnn, double dt, double* __restrict__ LO) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.N * m.L) return;
    const int n = (int)(t / m.L), k = (int)(t - (int64_t)n * m.L);
    const int L = m.L;
    const int2 nr = m.nlev[n];
    double acc = 0.0;
    for (int p = m.ecsr_off[n], pe = m.ecsr_off[n + 1]; p < pe; ++p) {
        const int ent = m.ecsr[p];
        const int e = ent >> 1;
        const int2 nd = m.edges[e];
        const double vf = vflux[(size_t)e * L + k];
        const double av = fabs(vf);
        const double T1 = T[(size_t)nd.x * L + k], T2 = T[(size_t)nd.y * L + k];
        const double f = -0.5 * (T1 * (vf + av) + T2 * (vf - av));
        acc = acc + ((ent & 1) ? -f : f);
    }
    const size_t i = (size_t)t;
    const double fv0 = flux_v_lo(m, T, w_e[i], n, k, nr);
    const double fv1 = (k + 1 < L) ? flux_v_lo(m, T, w_e[i + 1], n, k + 1, nr) : 0.0;
    const double div_v = fv0 - fv1;
    const double a = m.areasvol[i], hn = hnn[i];
    const double safe_a = a > 0.0 ? a : 1.0, safe_hn = hn > 0.0 ? hn : 1.0;
    const double numer = T[i] * hnode[i] + (acc + div_v) * dt / safe_a;
    const double lo = numer / safe_hn;
    LO[i] = (in_layer(nr, k) && hn > 0.0 && a > 0.0) ? lo : 0.0;
}

// adv_tra_vert_impl + ops.tdma (tracer_adv.py:487-529, ops.py:121-168); one thread
// per node column, updates X in place. cp/dp are [N][L] scratch.
__global__ void __launch_bounds__(256)
k_vert_impl(DevMesh m, const double* __restrict__ w_i, const double* __restrict__ hnn,
            double dt, double* __restrict__ X, double* __restrict__ cp,
            double* __restrict__ dp) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= m.N) return;
    const int L = m.L;
    const int2 nr = m.nlev[n];
    const size_t b0 = (size_t)n * L;
    double cp_prev = 0.0, dp_prev = 0.0;
    for (int k = 0; k < L; ++k) {
        const size_t i = b0 + k;
        const VimplRow r = vimpl_row(m, w_i, hnn, dt, b0, k, nr);
        const double Td = X[k > 0 ? i - 1 : i];
        const double Tc = X[i];
        const double Tu = k + 1 < L ? X[i + 1] : 0.0;
        const double rhs = r.valid ? -(r.a * Td + r.b_adv * Tc + r.c * Tu) : 0.0;
        const double mm = r.b - r.a * cp_prev;
        const double cpi = r.c / mm;
        const double dpi = (rhs - r.a * dp_prev) / mm;
        cp[i] = cpi; dp[i] = dpi;
        cp_prev = cpi; dp_prev = dpi;
    }
    double x_next = 0.0;
    for (int k = L - 1; k >= 0; --k) {
        const size_t i = b0 + k;
        const double x = dp[i] - cp[i] * x_next;
        x_next = x;
        cp[i] = x;   // keep x until the column is done (X[k-1] is still read below)
    }
    for (int k = 0; k < L; ++k) {
        const size_t i = b0 + k;
        X[i] = X[i] + (in_layer(nr, k) ? cp[i] : 0.0);
    }
}

// tracer_gradient_elements, tracer_adv.py:316-326; tr is [E][L][2].
__global__ void __launch_bounds__(256)
k_trxy(DevMesh m, const double* __restrict__ T, double2* __restrict__ tr) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.E * m.L) return;
    const int e = (int)(t / m.L), k = (int)(t - (int64_t)e * m.L);
    double2 g = make_double2(0.0, 0.0);
    if (in_layer(m.elev[e], k)) {
        const int L = m.L;
        const int v0 = m.en[3 * e], v1 = m.en[3 * e + 1], v2 = m.en[3 * e + 2];
        const double t0 = T[(size_t)v0 * L + k], t1 = T[(size_t)v1 * L + k],
                     t2 = T[(size_t)v2 * L + k];
        const double* gs = m.gs + 6 * (size_t)e;
        g.x = gs[0] * t0 + gs[1] * t1 + gs[2] * t2;
        g.y = gs[3] * t0 + gs[4] * t1 + gs[5] * t2;
    }
    tr[t] = g;
}

// _node_avg_grad, tracer_adv.py:329-341.
__global__ void __launch_bounds__(256)
k_gnode(DevMesh m, const double2* __restrict__ tr, double2* __restrict__ gnode) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.N * m.L) return;
    const int n = (int)(t / m.L), k = (int)(t - (int64_t)n * m.L);
    double nx = 0.0, ny = 0.0, den = 0.0;
    for (int p = m.ncsr_off[n], pe = m.ncsr_off[n + 1]; p < pe; ++p) {
        const int e = m.ncsr[p];
        const double aw = in_layer(m.elev[e], k) ? m.earea[e] : 0.0;
        const double2 g = tr[(size_t)e * m.L + k];
        nx = nx + g.x * aw;
        ny = ny + g.y * aw;
        den = den + aw;
    }
    const double d = den > 0.0 ? den : 1.0;
    gnode[t] = make_double2(nx / d, ny / d);
}

// adf_h = adv_flux_hor_ho(ttfAB, fill_up_dn_grad) - adv_flux_hor(T)
// (tracer_adv.py:196-225, 344-368, 564).
__global__ void __launch_bounds__(256)
k_adfh(DevMesh m, const double* __restrict__ vflux, const double* __restrict__ T,
       const double* __restrict__ Told, const double2* __restrict__ tr,
       const double2* __restrict__ gnode, double* __restrict__ adfh) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.D * m.L) return;
    const int e = (int)(t / m.L), k = (int)(t - (int64_t)e * m.L);
    const int L = m.L;
    const int2 nd = m.edges[e];
    const double vf = vflux[t];
    const double av = fabs(vf);
    const size_t i1 = (size_t)nd.x * L + k, i2 = (size_t)nd.y * L + k;
    const double T1 = T[i1], T2 = T[i2];
    const double flo = -0.5 * (T1 * (vf + av) + T2 * (vf - av));
    const double A1 = AB_OLD * Told[i1] + AB_NEW * T1;
    const double A2 = AB_OLD * Told[i2] + AB_NEW * T2;
    const double diff = A2 - A1;
    const int2 ur = m.utri_rng[e];
    double g0, g1, g2, g3;
    if (k >= ur.x && k < ur.y) {
        const int2 ud = m.updn[e];
        const double2 gu = tr[(size_t)ud.x * L + k], gd = tr[(size_t)ud.y * L + k];
        g0 = gu.x; g1 = gd.x; g2 = gu.y; g3 = gd.y;
    } else {
        const double2 ga = gnode[i1], gb = gnode[i2];
        g0 = ga.x; g1 = gb.x; g2 = ga.y; g3 = gb.y;
    }
    const double2 cc = m.edge_c[e];   // (edx*a, edy*R_EARTH)
    const double Tm1 = A1 + (2.0 * diff + cc.x * g0 + cc.y * g2) / 6.0;
    const double Tm2 = A2 - (2.0 * diff + cc.x * g1 + cc.y * g3) / 6.0;
    const double ho = -0.5 * ((vf + av) * Tm1 + (vf - av) * Tm2);
    adfh[t] = ho - flo;
}

// adf_v = adv_flux_ver_ho(ttfAB) - adv_flux_ver(T) (tracer_adv.py:241-295, 567-572);
// with w_i the full w = w_e + w_i is used in both terms.
__global__ void __launch_bounds__(256)
k_adfv(DevMesh m, const double* __restrict__ T, const double* __restrict__ Told,
       const double* __restrict__ w_e, const double* __restrict__ w_i,
       const double* __restrict__ Z3d, const double* __restrict__ zbar3,
       double* __restrict__ adfv) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.N * m.L) return;
    const int n = (int)(t / m.L), k = (int)(t - (int64_t)n * m.L);
    const int L = m.L;
    const int2 nr = m.nlev[n];
    const size_t i = (size_t)t;
    const double W = w_at(w_e, w_i, i);
    double hi = 0.0;
    if (in_iface(nr, k)) {
        const size_t b0 = (size_t)n * L;
        const int km1 = k > 0 ? k - 1 : 0, km2 = k > 1 ? k - 2 : 0;
        const double ar = m.area[i];
        const double A = AB_OLD * Told[i] + AB_NEW * T[i];
        const double Am1 = AB_OLD * Told[b0 + km1] + AB_NEW * T[b0 + km1];
        if (k == nr.x) {
            hi = -A * W * ar;
        } else if (k == nr.y) {
            hi = 0.0;
        } else if (k == nr.x + 1 || k == nr.y - 1) {
            hi = -0.5 * (Am1 + A) * W * ar;
        } else if (k >= nr.x + 2 && k <= nr.y - 2) {
            const double Am2 = AB_OLD * Told[b0 + km2] + AB_NEW * T[b0 + km2];
            const double Ap1 = (k + 1 < L) ? AB_OLD * Told[i + 1] + AB_NEW * T[i + 1] : 0.0;
            double Znz, Zm1, Zm2, Zp1, zb;
            if (Z3d) {
                const int kp1 = k + 1 < L ? k + 1 : L - 1;
                Znz = Z3d[i]; Zm1 = Z3d[b0 + km1]; Zm2 = Z3d[b0 + km2]; Zp1 = Z3d[b0 + kp1];
                zb = zbar3[i];
            } else {
                Znz = m.Zs[k]; Zm1 = m.Zs[L + k]; Zm2 = m.Zs[2 * L + k]; Zp1 = m.Zs[3 * L + k];
                zb = m.zbar[k];
            }
            const double qc = (Am1 - A) / nz_safe(Zm1 - Znz);
            const double qu = (A - Ap1) / nz_safe(Znz - Zp1);
            const double qd = (Am2 - Am1) / nz_safe(Zm2 - Zm1);
            const double Tmean1 = A + (2.0 * qc + qu) * (zb - Znz) / 3.0;
            const double Tmean2 = Am1 + (2.0 * qc + qd) * (zb - Zm1) / 3.0;
            hi = -0.5 * (Tmean1 + Tmean2) * W * ar;
        }
    }
    const double lo = flux_v_lo(m, T, W, n, k, nr);
    adfv[i] = hi - lo;
}

// zalesak_limit a1-a3 (tracer_adv.py:389-405): per node, the max/min over its
// elements of the element's 3-vertex max/min of max/min(LO, T); -/+1e3 for an
// element that is dry at k; -/+inf for a node without elements.
__global__ void __launch_bounds__(256)
k_bounds(DevMesh m, const double* __restrict__ LO, const double* __restrict__ T,
         double* __restrict__ tvmax, double* __restrict__ tvmin) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.N * m.L) return;
    const int n = (int)(t / m.L), k = (int)(t - (int64_t)n * m.L);
    const int L = m.L;
    double mx = -INFINITY, mn = INFINITY;
    for (int p = m.ncsr_off[n], pe = m.ncsr_off[n + 1]; p < pe; ++p) {
        const int e = m.ncsr[p];
        double amax = -FCT_BIG, amin = FCT_BIG;
        if (in_layer(m.elev[e], k)) {
            const size_t j0 = (size_t)m.en[3 * e] * L + k, j1 = (size_t)m.en[3 * e + 1] * L + k,
                         j2 = (size_t)m.en[3 * e + 2] * L + k;
            const double l0 = LO[j0], l1 = LO[j1], l2 = LO[j2];
            const double t0 = T[j0], t1 = T[j1], t2 = T[j2];
            amax = fmax(fmax(fmax(l0, t0), fmax(l1, t1)), fmax(l2, t2));
            amin = fmin(fmin(fmin(l0, t0), fmin(l1, t1)), fmin(l2, t2));
        }
        mx = fmax(mx, amax);
        mn = fmin(mn, amin);
    }
    tvmax[t] = mx;
    tvmin[t] = mn;
}

// zalesak_limit b1 (tracer_adv.py:426-435): fct_plus / fct_minus.
__global__ void __launch_bounds__(256)
k_pm(DevMesh m, const double* __restrict__ adfh, const double* __restrict__ adfv,
     double* __restrict__ fplus, double* __restrict__ fminus) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.N * m.L) return;
    const int n = (int)(t / m.L), k = (int)(t - (int64_t)n * m.L);
    const int L = m.L;
    const int2 nr = m.nlev[n];
    const size_t i = (size_t)t;
    const double fv = adfv[i];
    const double fb = (k + 1 < L) ? adfv[i + 1] : 0.0;
    double pos_v = fmax(fv, 0.0) + fmax(-fb, 0.0);
    double neg_v = fmin(fv, 0.0) + fmin(-fb, 0.0);
    if (!in_layer(nr, k)) { pos_v = 0.0; neg_v = 0.0; }
    double ap = 0.0, am = 0.0;
    for (int p = m.ecsr_off[n], pe = m.ecsr_off[n + 1]; p < pe; ++p) {
        const int ent = m.ecsr[p];
        const double f0 = adfh[(size_t)(ent >> 1) * L + k];
        const double f = (ent & 1) ? -f0 : f0;
        ap = ap + fmax(f, 0.0);
        am = am + fmin(f, 0.0);
    }
    fplus[i] = pos_v + ap;
    fminus[i] = neg_v + am;
}

// zalesak_limit a4 + b2 (tracer_adv.py:407-424, 441-451): limiter factors.
__global__ void __launch_bounds__(256)
k_fac(DevMesh m, const double* __restrict__ LO, const double* __restrict__ tvmax,
      const double* __restrict__ tvmin, const double* __restrict__ fplus,
      const double* __restrict__ fminus, const double* __restrict__ hnn, double dt,
      double* __restrict__ pf, double* __restrict__ mf) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.N * m.L) return;
    const int n = (int)(t / m.L), k = (int)(t - (int64_t)n * m.L);
    const int L = m.L;
    const int2 nr = m.nlev[n];
    const size_t i = (size_t)t;
    const size_t im = k > 0 ? i - 1 : i;
    const bool use_own = (k == nr.x) || (k == nr.y - 1);
    const bool nlm = in_layer(nr, k);
    const double lo = LO[i];
    double cmax, cmin;
    if (use_own) {
        cmax = tvmax[i];
        cmin = tvmin[i];
    } else {
        cmax = fmax(fmax(tvmax[im], tvmax[i]), (k + 1 < L) ? tvmax[i + 1] : 0.0);
        cmin = fmin(fmin(tvmin[im], tvmin[i]), (k + 1 < L) ? tvmin[i + 1] : 0.0);
    }
    const double ttf_max = nlm ? cmax - lo : 0.0;
    const double ttf_min = nlm ? cmin - lo : 0.0;
    const double a = m.areasvol[i], hn = hnn[i];
    const double safe_a = a > 0.0 ? a : 1.0, safe_hn = hn > 0.0 ? hn : 1.0;
    const double flux_pos = fplus[i] * dt / safe_a / safe_hn + FCT_FLUX_EPS;
    const double flux_neg = fminus[i] * dt / safe_a / safe_hn - FCT_FLUX_EPS;
    const bool valid = nlm && a > 0.0 && hn > 0.0;
    pf[i] = valid ? fmin(1.0, ttf_max / flux_pos) : 1.0;
    mf[i] = valid ? fmin(1.0, ttf_min / flux_neg) : 1.0;
}

// Limited vertical antidiffusive flux at interface kk (tracer_adv.py:453-459).
__device__ __forceinline__ double adfv_lim(const DevMesh& m, const double* __restrict__ adfv,
                                           const double* __restrict__ pf,
                                           const double* __restrict__ mf, size_t b0,
                                           int kk, int2 nr) {
    if (!in_iface(nr, kk)) return 0.0;
    const double f = adfv[b0 + kk];
    double ae;
    if (kk == nr.x) {
        ae = f >= 0.0 ? pf[b0 + kk] : mf[b0 + kk];
    } else {
        const size_t km = b0 + (kk > 0 ? kk - 1 : 0);
        ae = f >= 0.0 ? fmin(mf[km], pf[b0 + kk]) : fmin(pf[km], mf[b0 + kk]);
    }
    return ae * f;
}

// zalesak_limit b3 + flux2dtracer_fct + ale_reconstruct (tracer_adv.py:453-466,
// 469-484, 142-151). Writes every lane of T_new (dry lanes: T + 0).
__global__ void __launch_bounds__(256)
k_final(DevMesh m, const double* __restrict__ adfh, const double* __restrict__ adfv,
        const double* __restrict__ pf, const double* __restrict__ mf,
        const double* __restrict__ T, const double* __restrict__ LO,
        const double* __restrict__ hnode, const double* __restrict__ hnn, double dt,
        double* __restrict__ Tnew) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.N * m.L) return;
    const int n = (int)(t / m.L), k = (int)(t - (int64_t)n * m.L);
    const int L = m.L;
    const int2 nr = m.nlev[n];
    const size_t i = (size_t)t;
    double acc = 0.0;
    for (int p = m.ecsr_off[n], pe = m.ecsr_off[n + 1]; p < pe; ++p) {
        const int ent = m.ecsr[p];
        const int e = ent >> 1;
        const int2 nd = m.edges[e];
        const double f = adfh[(size_t)e * L + k];
        const size_t i1 = (size_t)nd.x * L + k, i2 = (size_t)nd.y * L + k;
        const double ae = f >= 0.0 ? fmin(pf[i1], mf[i2]) : fmin(mf[i1], pf[i2]);
        const double fl = ae * f;
        acc = acc + ((ent & 1) ? -fl : fl);
    }
    const size_t b0 = (size_t)n * L;
    const double vk = adfv_lim(m, adfv, pf, mf, b0, k, nr);
    const double vk1 = (k + 1 < L) ? adfv_lim(m, adfv, pf, mf, b0, k + 1, nr) : 0.0;
    const double a = m.areasvol[i], hn = hnn[i];
    const double safe_a = a > 0.0 ? a : 1.0, safe_hn = hn > 0.0 ? hn : 1.0;
    const bool nlm = in_layer(nr, k);
    const double Ti = T[i];
    const double lo_trans = -Ti * hnode[i] + LO[i] * hn;
    const double vdiv = a > 0.0 ? (vk - vk1) * dt / safe_a : 0.0;
    const double dttf_v = nlm ? lo_trans + vdiv : 0.0;
    const double dttf_h = nlm ? acc * dt / safe_a : 0.0;
    double del = dttf_h + dttf_v;
    del = del + Ti * (hnode[i] - hn);
    const double dT = (nlm && hn > 0.0) ? del / safe_hn : 0.0;
    Tnew[i] = Ti + dT;
}

// ---------------------------------------------------------------------------
// single stages with the reference's inputs and outputs (fct_stage): each kernel
// restates one public function of fesom_jax/tracer_adv.py on every lane of its output,
// in the expression order of the stage kernels above
// ---------------------------------------------------------------------------
// tracer_ab, tracer_adv.py:59-66 (ttfAB only; T_old_new is the input T)
__global__ void __launch_bounds__(256)
k_st_ab(int64_t count, const double* __restrict__ T, const double* __restrict__ Told,
        double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) out[t] = AB_OLD * Told[t] + AB_NEW * T[t];
}

// adv_flux_hor, tracer_adv.py:96-105: upwind flux of ttf through each edge-level
__global__ void __launch_bounds__(256)
k_st_hlo(DevMesh m, const double* __restrict__ vflux, const double* __restrict__ ttf,
         double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.D * m.L) return;
    const int e = (int)(t / m.L), k = (int)(t - (int64_t)e * m.L);
    const int2 nd = m.edges[e];
    const double vf = vflux[t];
    const double av = fabs(vf);
    const double T1 = ttf[(size_t)nd.x * m.L + k], T2 = ttf[(size_t)nd.y * m.L + k];
    out[t] = -0.5 * (T1 * (vf + av) + T2 * (vf - av));
}

// adv_flux_ver, tracer_adv.py:108-120
__global__ void __launch_bounds__(256)
k_st_vlo(DevMesh m, const double* __restrict__ w_e, const double* __restrict__ ttf,
         double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.N * m.L) return;
    const int n = (int)(t / m.L), k = (int)(t - (int64_t)n * m.L);
    out[t] = flux_v_lo(m, ttf, w_e[t], n, k, m.nlev[n]);
}

// compute_fct_lo, tracer_adv.py:298-313, from given edge and vertical fluxes
__global__ void __launch_bounds__(256)
k_st_lo(DevMesh m, const double* __restrict__ fh, const double* __restrict__ fv,
        const double* __restrict__ T, const double* __restrict__ hnode,
        const double* __restrict__ hnn, double dt, double* __restrict__ LO) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.N * m.L) return;
    const int n = (int)(t / m.L), k = (int)(t - (int64_t)n * m.L);
    const int L = m.L;
    double acc = 0.0;
    for (int p = m.ecsr_off[n], pe = m.ecsr_off[n + 1]; p < pe; ++p) {
        const int ent = m.ecsr[p];
        const double f = fh[(size_t)(ent >> 1) * L + k];
        acc = acc + ((ent & 1) ? -f : f);
    }
    const size_t i = (size_t)t;
    const double div_v = fv[i] - ((k + 1 < L) ? fv[i + 1] : 0.0);
    const double a = m.areasvol[i], hn = hnn[i];
    const double safe_a = a > 0.0 ? a : 1.0, safe_hn = hn > 0.0 ? hn : 1.0;
    const double numer = T[i] * hnode[i] + (acc + div_v) * dt / safe_a;
    const double lo = numer / safe_hn;
    LO[i] = (in_layer(m.nlev[n], k) && hn > 0.0 && a > 0.0) ? lo : 0.0;
}

// fill_up_dn_grad, tracer_adv.py:344-368: [D][L][4] gradient slots from tr_xy and the
// node averages
__global__ void __launch_bounds__(256)
k_st_eud(DevMesh m, const double2* __restrict__ tr, const double2* __restrict__ gnode,
         double4* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.D * m.L) return;
    const int e = (int)(t / m.L), k = (int)(t - (int64_t)e * m.L);
    const int L = m.L;
    const int2 ur = m.utri_rng[e];
    if (k >= ur.x && k < ur.y) {
        const int2 ud = m.updn[e];
        const double2 gu = tr[(size_t)ud.x * L + k], gd = tr[(size_t)ud.y * L + k];
        out[t] = make_double4(gu.x, gd.x, gu.y, gd.y);
    } else {
        const int2 nd = m.edges[e];
        const double2 ga = gnode[(size_t)nd.x * L + k], gb = gnode[(size_t)nd.y * L + k];
        out[t] = make_double4(ga.x, gb.x, ga.y, gb.y);
    }
}

// adv_flux_hor_ho, tracer_adv.py:196-225, from given gradient slots
__global__ void __launch_bounds__(256)
k_st_hho(DevMesh m, const double* __restrict__ vflux, const double* __restrict__ ttf,
         const double4* __restrict__ eud, double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.D * m.L) return;
    const int e = (int)(t / m.L), k = (int)(t - (int64_t)e * m.L);
    const int2 nd = m.edges[e];
    const double T1 = ttf[(size_t)nd.x * m.L + k], T2 = ttf[(size_t)nd.y * m.L + k];
    const double diff = T2 - T1;
    const double4 g = eud[t];
    const double2 cc = m.edge_c[e];
    const double Tm1 = T1 + (2.0 * diff + cc.x * g.x + cc.y * g.z) / 6.0;
    const double Tm2 = T2 - (2.0 * diff + cc.x * g.y + cc.y * g.w) / 6.0;
    const double vf = vflux[t];
    const double av = fabs(vf);
    out[t] = -0.5 * ((vf + av) * Tm1 + (vf - av) * Tm2);
}

// adv_flux_ver_ho, tracer_adv.py:241-295 (static stencil, or Z3d / zbar3)
__global__ void __launch_bounds__(256)
k_st_vho(DevMesh m, const double* __restrict__ ttf, const double* __restrict__ w_e,
         const double* __restrict__ Z3d, const double* __restrict__ zbar3,
         double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.N * m.L) return;
    const int n = (int)(t / m.L), k = (int)(t - (int64_t)n * m.L);
    const int L = m.L;
    const int2 nr = m.nlev[n];
    const size_t i = (size_t)t;
    double hi = 0.0;
    if (in_iface(nr, k)) {
        const double W = w_e[i];
        const size_t b0 = (size_t)n * L;
        const int km1 = k > 0 ? k - 1 : 0, km2 = k > 1 ? k - 2 : 0;
        const double ar = m.area[i];
        const double A = ttf[i], Am1 = ttf[b0 + km1];
        if (k == nr.x) {
            hi = -A * W * ar;
        } else if (k == nr.y) {
            hi = 0.0;
        } else if (k == nr.x + 1 || k == nr.y - 1) {
            hi = -0.5 * (Am1 + A) * W * ar;
        } else if (k >= nr.x + 2 && k <= nr.y - 2) {
            const double Am2 = ttf[b0 + km2];
            const double Ap1 = (k + 1 < L) ? ttf[i + 1] : 0.0;
            double Znz, Zm1, Zm2, Zp1, zb;
            if (Z3d) {
                const int kp1 = k + 1 < L ? k + 1 : L - 1;
                Znz = Z3d[i]; Zm1 = Z3d[b0 + km1]; Zm2 = Z3d[b0 + km2]; Zp1 = Z3d[b0 + kp1];
                zb = zbar3[i];
            } else {
                Znz = m.Zs[k]; Zm1 = m.Zs[L + k]; Zm2 = m.Zs[2 * L + k]; Zp1 = m.Zs[3 * L + k];
                zb = m.zbar[k];
            }
            const double qc = (Am1 - A) / nz_safe(Zm1 - Znz);
            const double qu = (A - Ap1) / nz_safe(Znz - Zp1);
            const double qd = (Am2 - Am1) / nz_safe(Zm2 - Zm1);
            const double Tmean1 = A + (2.0 * qc + qu) * (zb - Znz) / 3.0;
            const double Tmean2 = Am1 + (2.0 * qc + qd) * (zb - Zm1) / 3.0;
            hi = -0.5 * (Tmean1 + Tmean2) * W * ar;
        }
    }
    out[i] = hi;
}

// zalesak_limit b3, tracer_adv.py:453-466: limited edge and vertical fluxes
__global__ void __launch_bounds__(256)
k_st_hlim(DevMesh m, const double* __restrict__ adfh, const double* __restrict__ pf,
          const double* __restrict__ mf, double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.D * m.L) return;
    const int e = (int)(t / m.L), k = (int)(t - (int64_t)e * m.L);
    const int2 nd = m.edges[e];
    const size_t i1 = (size_t)nd.x * m.L + k, i2 = (size_t)nd.y * m.L + k;
    const double f = adfh[t];
    const double ae = f >= 0.0 ? fmin(pf[i1], mf[i2]) : fmin(mf[i1], pf[i2]);
    out[t] = ae * f;
}

__global__ void __launch_bounds__(256)
k_st_vlim(DevMesh m, const double* __restrict__ adfv, const double* __restrict__ pf,
          const double* __restrict__ mf, double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.N * m.L) return;
    const int n = (int)(t / m.L), k = (int)(t - (int64_t)n * m.L);
    out[t] = adfv_lim(m, adfv, pf, mf, (size_t)n * m.L, k, m.nlev[n]);
}

// flux2dtracer_fct, tracer_adv.py:469-484 (from limited fluxes)
__global__ void __launch_bounds__(256)
k_st_f2d(DevMesh m, const double* __restrict__ T, const double* __restrict__ LO,
         const double* __restrict__ adfh, const double* __restrict__ adfv,
         const double* __restrict__ hnode, const double* __restrict__ hnn, double dt,
         double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.N * m.L) return;
    const int n = (int)(t / m.L), k = (int)(t - (int64_t)n * m.L);
    const int L = m.L;
    double acc = 0.0;
    for (int p = m.ecsr_off[n], pe = m.ecsr_off[n + 1]; p < pe; ++p) {
        const int ent = m.ecsr[p];
        const double f = adfh[(size_t)(ent >> 1) * L + k];
        acc = acc + ((ent & 1) ? -f : f);
    }
    const size_t i = (size_t)t;
    const double a = m.areasvol[i];
    const double safe_a = a > 0.0 ? a : 1.0;
    const bool nlm = in_layer(m.nlev[n], k);
    const double lo_trans = -T[i] * hnode[i] + LO[i] * hnn[i];
    const double vdiv = a > 0.0 ? (adfv[i] - ((k + 1 < L) ? adfv[i + 1] : 0.0)) * dt / safe_a : 0.0;
    const double dttf_v = nlm ? lo_trans + vdiv : 0.0;
    const double dttf_h = nlm ? acc * dt / safe_a : 0.0;
    out[i] = dttf_h + dttf_v;
}

// flux2dtracer, tracer_adv.py:123-139 (the upwind driver's assembly)
__global__ void __launch_bounds__(256)
k_st_f2d_up(DevMesh m, const double* __restrict__ fh, const double* __restrict__ fv, double dt,
            double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.N * m.L) return;
    const int n = (int)(t / m.L), k = (int)(t - (int64_t)n * m.L);
    const int L = m.L;
    double acc = 0.0;
    for (int p = m.ecsr_off[n], pe = m.ecsr_off[n + 1]; p < pe; ++p) {
        const int ent = m.ecsr[p];
        const double f = fh[(size_t)(ent >> 1) * L + k];
        acc = acc + ((ent & 1) ? -f : f);
    }
    const size_t i = (size_t)t;
    const double a = m.areasvol[i];
    const double safe = a > 0.0 ? a : 1.0;
    const bool lm = in_layer(m.nlev[n], k);
    const double dv = lm ? (fv[i] - ((k + 1 < L) ? fv[i + 1] : 0.0)) * dt / safe : 0.0;
    const double dh = lm ? acc * dt / safe : 0.0;
    out[i] = dh + dv;
}

// ale_reconstruct, tracer_adv.py:142-151
__global__ void __launch_bounds__(256)
k_st_ale(DevMesh m, const double* __restrict__ T, const double* __restrict__ del,
         const double* __restrict__ hnode, const double* __restrict__ hnn,
         double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.N * m.L) return;
    const int n = (int)(t / m.L), k = (int)(t - (int64_t)n * m.L);
    const size_t i = (size_t)t;
    const double Ti = T[i], hn = hnn[i];
    const double d = del[i] + Ti * (hnode[i] - hn);
    const double safe_hn = hn > 0.0 ? hn : 1.0;
    const double dT = (in_layer(m.nlev[n], k) && hn > 0.0) ? d / safe_hn : 0.0;
    out[i] = Ti + dT;
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
static inline unsigned grid_for(int64_t count, int block) {
    return (unsigned)((count + block - 1) / block);
}

#define FCT_LAUNCH(kern, count, stream, ...)                                          \
    do {                                                                              \
        const int64_t c_ = (count);                                                   \
        if (c_ > 0) hipLaunchKernelGGL(kern, dim3(grid_for(c_, 256)), dim3(256), 0,  \
                                       (stream), __VA_ARGS__);                        \
    } while (0)

int launch_vflux(const DevMesh& m, const fct_step_args* a, double* vflux, hipStream_t s) {
    FCT_LAUNCH(k_vflux, (int64_t)m.D * m.L, s, m, a->uv, a->helem, vflux);
    return check_launch("k_vflux");
}

int launch_lo(const DevMesh& m, const fct_step_args* a, int j, const double* vflux,
              double* LO, double* scr0, double* scr1, hipStream_t s) {
    FCT_LAUNCH(k_lo, (int64_t)m.N * m.L, s, m, vflux, a->T[j], a->w_e, a->hnode,
               a->hnode_new, a->dt, LO);
    int rc = check_launch("k_lo");
    if (rc) return rc;
    if (a->w_i) {
        FCT_LAUNCH(k_vert_impl, (int64_t)m.N, s, m, a->w_i, a->hnode_new, a->dt, LO, scr0, scr1);
        rc = check_launch("k_vert_impl");
    }
    return rc;
}

int launch_vert_impl(const DevMesh& m, const double* w_i, const double* hnn, double dt,
                     double* X, double* cp, double* dp, hipStream_t s) {
    FCT_LAUNCH(k_vert_impl, (int64_t)m.N, s, m, w_i, hnn, dt, X, cp, dp);
    return check_launch("k_vert_impl");
}

int launch_grad(const DevMesh& m, const fct_step_args* a, int j, double* tr, hipStream_t s) {
    FCT_LAUNCH(k_trxy, (int64_t)m.E * m.L, s, m, a->T[j], (double2*)tr);
    return check_launch("k_trxy");
}

int launch_pm(const DevMesh& m, const fct_step_args* a, int j, const double* vflux,
              const double* LO, const double* tr, double* gnode, double* adfh, double* adfv,
              double* tvmax, double* tvmin, double* fplus, double* fminus, hipStream_t s) {
    const int64_t nl = (int64_t)m.N * m.L, dl = (int64_t)m.D * m.L;
    FCT_LAUNCH(k_gnode, nl, s, m, (const double2*)tr, (double2*)gnode);
    FCT_LAUNCH(k_adfh, dl, s, m, vflux, a->T[j], a->T_old[j], (const double2*)tr,
               (const double2*)gnode, adfh);
    FCT_LAUNCH(k_adfv, nl, s, m, a->T[j], a->T_old[j], a->w_e, a->w_i, a->Z3d, a->zbar3, adfv);
    FCT_LAUNCH(k_bounds, nl, s, m, LO, a->T[j], tvmax, tvmin);
    FCT_LAUNCH(k_pm, nl, s, m, adfh, adfv, fplus, fminus);
    return check_launch("phase 3 kernels");
}

int launch_fac(const DevMesh& m, const fct_step_args* a, const double* LO, const double* tvmax,
               const double* tvmin, const double* fplus, const double* fminus, double* pf,
               double* mf, hipStream_t s) {
    FCT_LAUNCH(k_fac, (int64_t)m.N * m.L, s, m, LO, tvmax, tvmin, fplus, fminus, a->hnode_new,
               a->dt, pf, mf);
    return check_launch("k_fac");
}

int launch_final(const DevMesh& m, const fct_step_args* a, int j, const double* LO,
                 const double* adfh, const double* adfv, const double* tvmax,
                 const double* tvmin, const double* fplus, const double* fminus, double* pf,
                 double* mf, hipStream_t s) {
    const int64_t nl = (int64_t)m.N * m.L;
    int rc = launch_fac(m, a, LO, tvmax, tvmin, fplus, fminus, pf, mf, s);
    if (rc) return rc;
    FCT_LAUNCH(k_final, nl, s, m, adfh, adfv, pf, mf, a->T[j], LO, a->hnode, a->hnode_new,
               a->dt, a->T_new[j]);
    return check_launch("phase 4 kernels");
}

// One reference stage (fct_stage): `in` and `out` as FCT_STAGE_* documents; scratch
// holds vflux [D][L], then (for the stages that need them) gnode [N][L][2] and six
// [N][L] arrays.
int launch_stage(const DevMesh& m, const StageScratch& w, int stage, const double* const* in,
                 double* const* out, double dt, hipStream_t s) {
    const int64_t nl = (int64_t)m.N * m.L, dl = (int64_t)m.D * m.L, el = (int64_t)m.E * m.L;
    switch (stage) {
        case FCT_STAGE_TRACER_AB:
            FCT_LAUNCH(k_st_ab, nl, s, nl, in[0], in[1], out[0]);
            break;
        case FCT_STAGE_EDGE_VFLUX:
            FCT_LAUNCH(k_vflux, dl, s, m, in[0], in[1], out[0]);
            break;
        case FCT_STAGE_ADV_FLUX_HOR:
            FCT_LAUNCH(k_vflux, dl, s, m, in[0], in[1], w.vflux);
            FCT_LAUNCH(k_st_hlo, dl, s, m, w.vflux, in[2], out[0]);
            break;
        case FCT_STAGE_ADV_FLUX_VER:
            FCT_LAUNCH(k_st_vlo, nl, s, m, in[0], in[1], out[0]);
            break;
        case FCT_STAGE_FLUX2DTRACER:
            FCT_LAUNCH(k_st_f2d_up, nl, s, m, in[0], in[1], dt, out[0]);
            break;
        case FCT_STAGE_ALE_RECONSTRUCT:
            FCT_LAUNCH(k_st_ale, nl, s, m, in[0], in[1], in[2], in[3], out[0]);
            break;
        case FCT_STAGE_COMPUTE_FCT_LO:
            FCT_LAUNCH(k_st_lo, nl, s, m, in[0], in[1], in[2], in[3], in[4], dt, out[0]);
            break;
        case FCT_STAGE_TRACER_GRADIENT_ELEMENTS:
            FCT_LAUNCH(k_trxy, el, s, m, in[0], (double2*)out[0]);
            break;
        case FCT_STAGE_FILL_UP_DN_GRAD:
            FCT_LAUNCH(k_gnode, nl, s, m, (const double2*)in[0], (double2*)w.gnode);
            FCT_LAUNCH(k_st_eud, dl, s, m, (const double2*)in[0], (const double2*)w.gnode,
                       (double4*)out[0]);
            break;
        case FCT_STAGE_ADV_FLUX_HOR_HO:
            FCT_LAUNCH(k_vflux, dl, s, m, in[0], in[1], w.vflux);
            FCT_LAUNCH(k_st_hho, dl, s, m, w.vflux, in[2], (const double4*)in[3], out[0]);
            break;
        case FCT_STAGE_ADV_FLUX_VER_HO:
            FCT_LAUNCH(k_st_vho, nl, s, m, in[1], in[0], in[2], in[3], out[0]);
            break;
        case FCT_STAGE_ZALESAK_PM:
            FCT_LAUNCH(k_pm, nl, s, m, in[2], in[3], out[0], out[1]);
            break;
        case FCT_STAGE_ZALESAK_LIMIT:
        case FCT_STAGE_ZALESAK_APPLY: {
            const double* fp = w.fplus;
            const double* fm = w.fminus;
            if (stage == FCT_STAGE_ZALESAK_LIMIT) {
                FCT_LAUNCH(k_pm, nl, s, m, in[2], in[3], w.fplus, w.fminus);
            } else {
                fp = in[5];
                fm = in[6];
            }
            FCT_LAUNCH(k_bounds, nl, s, m, in[1], in[0], w.tvmax, w.tvmin);
            FCT_LAUNCH(k_fac, nl, s, m, in[1], w.tvmax, w.tvmin, fp, fm, in[4], dt, w.pf, w.mf);
            FCT_LAUNCH(k_st_hlim, dl, s, m, in[2], w.pf, w.mf, out[0]);
            FCT_LAUNCH(k_st_vlim, nl, s, m, in[3], w.pf, w.mf, out[1]);
            break;
        }
        case FCT_STAGE_FLUX2DTRACER_FCT:
            FCT_LAUNCH(k_st_f2d, nl, s, m, in[0], in[1], in[2], in[3], in[4], in[5], dt, out[0]);
            break;
        case FCT_STAGE_ADV_TRA_VERT_IMPL: {
            if (nl > 0 && hipMemcpyAsync(out[0], in[1], (size_t)nl * sizeof(double),
                                         hipMemcpyDeviceToDevice, s) != hipSuccess)
                return set_error(FCT_E_HIP, "adv_tra_vert_impl: copy of T failed");
            FCT_LAUNCH(k_vert_impl, (int64_t)m.N, s, m, in[0], in[2], dt, out[0], w.pf, w.mf);
            break;
        }
        default:
            return set_error(FCT_E_INVALID, "unknown stage %d", stage);
    }
    return check_launch("stage kernels");
}

int launch_tracer_ab(const double* T, const double* Told, double* out, int64_t count,
                     hipStream_t s) {
    FCT_LAUNCH(k_st_ab, count, s, count, T, Told, out);
    return check_launch("k_st_ab");
}

}  // namespace fct
