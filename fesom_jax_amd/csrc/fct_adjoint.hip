// Reverse mode of one FP64 FCT step for gfx950 (with or without the implicit split
// w_i and the moving-level stencil Z3d/zbar3, which are treated as constants).
//
// The forward tape is the phased path's dense intermediates of the same step (vflux,
// LO, tr_xy, gnode, adf_h, adf_v, the node bounds tvmax/tvmin, P+/P-, the limiter
// factors), recomputed for every step rather than stored. The kernels below then run
// the chain rule backwards through fesom_jax/tracer_adv.py:532-578 with the
// conventions of jax.grad on the reference (a plain subgradient,
// fesom_jax/docs/LIMITER_GRADIENTS.md:45-75):
//   max/min(x, y): the cotangent goes to the larger/smaller argument, half each on a tie;
//   a max/min over the 3 vertices or over a node's elements: shared equally by the
//   arguments equal to the result; abs'(0) = +1; where(): the chosen branch only;
//   x / y: g / y to x and (-g * x) * (1 / (y * y)) to y.
// Every transpose of a scatter is a gather over a CSR list (node -> (edge, side),
// node -> element, element -> (edge, side), element -> (edge, up/down role)), so the
// result is deterministic. Every kernel runs one thread per (entity, level); terms that
// couple levels are gathered by the receiving level (no thread writes another's lane).

#include "fct_common.h"

namespace fct {

namespace {

__device__ __forceinline__ double ab2a(double told, double t) { return AB_OLD * told + AB_NEW * t; }

// cotangents of max(x, y) / min(x, y) for x
__device__ __forceinline__ double gmax_x(double g, double x, double y) {
    return x > y ? g : (x == y ? 0.5 * g : 0.0);
}
__device__ __forceinline__ double gmin_x(double g, double x, double y) {
    return x < y ? g : (x == y ? 0.5 * g : 0.0);
}
__device__ __forceinline__ double gabs(double g, double x) { return x >= 0.0 ? g : -g; }

#define FLAT_NL(m)                                                               \
    const int64_t t_ = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;          \
    if (t_ >= (int64_t)(m).N * (m).L) return;                                   \
    const int n = (int)(t_ / (m).L), k = (int)(t_ - (int64_t)n * (m).L);        \
    const size_t i = (size_t)t_;                                                \
    (void)k; (void)i

// ---- flux2dtracer_fct + ale_reconstruct (tracer_adv.py:469-484, 142-151) --------
__global__ void __launch_bounds__(256)
ka_final(DevMesh m, const double* __restrict__ g_out, const double* __restrict__ hnode,
         const double* __restrict__ hnn, double dt, double* __restrict__ gT,
         double* __restrict__ gLO, double* __restrict__ gx, double* __restrict__ ghsc) {
    FLAT_NL(m);
    const int2 nr = m.nlev[n];
    const bool nlm = in_layer(nr, k);
    const double a = m.areasvol[i], hn = hnn[i];
    const double safe_a = a > 0.0 ? a : 1.0, safe_hn = hn > 0.0 ? hn : 1.0;
    const double g = g_out[i];
    const double g_del = (nlm && hn > 0.0) ? g / safe_hn : 0.0;
    double t = g + g_del * (hnode[i] - hn);
    const double g_in = nlm ? g_del : 0.0;
    t = t + (-(g_in * hnode[i]));
    gT[i] = t;
    gLO[i] = g_in * hn;
    gx[i] = (a > 0.0 ? g_in : 0.0) / safe_a * dt;
    ghsc[i] = g_in / safe_a * dt;
}

// ---- limited vertical flux (tracer_adv.py:453-459) -----------------------------------
// cotangent of the limited flux at interface kk (g_vl = gx[kk] - gx[kk-1]), its factor
// ae and g_vl * adf_v
__device__ __forceinline__ void vlim_at(const double* gx, const double* adfv, const double* pf,
                                        const double* mf, size_t b0, int kk, int2 nr,
                                        double& gvl, double& ae, double& g_aev, bool& pv) {
    const size_t i = b0 + kk, im = b0 + (kk > 0 ? kk - 1 : 0);
    const double gv = kk > 0 ? gx[i] + (-gx[im]) : gx[i];
    gvl = in_iface(nr, kk) ? gv : 0.0;
    const double f = adfv[i];
    pv = f >= 0.0;
    if (kk == nr.x) ae = pv ? pf[i] : mf[i];
    else ae = pv ? fmin(mf[im], pf[i]) : fmin(pf[im], mf[i]);
    g_aev = gvl * f;
}

// g_adfv (limited-flux part) and the factors' cotangents at level k, which gather the
// terms of interface k (its own factors) and of interface k+1 (min(mf[k], pf[k+1]) or
// min(pf[k], mf[k+1])); interface 0's "k-1" factor is level 0 itself.
__global__ void __launch_bounds__(256)
ka_vlim(DevMesh m, const double* __restrict__ gx, const double* __restrict__ adfv,
        const double* __restrict__ pf, const double* __restrict__ mf,
        double* __restrict__ g_adfv, double* __restrict__ g_pf, double* __restrict__ g_mf) {
    FLAT_NL(m);
    const int L = m.L;
    const int2 nr = m.nlev[n];
    const size_t b0 = (size_t)n * L, im = b0 + (k > 0 ? k - 1 : 0);
    double gvl, ae, g_aev;
    bool pv;
    vlim_at(gx, adfv, pf, mf, b0, k, nr, gvl, ae, g_aev, pv);
    g_adfv[i] = gvl * ae;
    double gp = 0.0, gm = 0.0;
    if (k == nr.x) {
        if (pv) gp += g_aev; else gm += g_aev;
    } else if (pv) {
        gp += gmin_x(g_aev, pf[i], mf[im]);
        if (k == 0) gm += gmin_x(g_aev, mf[im], pf[i]);
    } else {
        gm += gmin_x(g_aev, mf[i], pf[im]);
        if (k == 0) gp += gmin_x(g_aev, pf[im], mf[i]);
    }
    if (k + 1 < L && k + 1 != nr.x) {
        double gvl1, ae1, g1;
        bool pv1;
        vlim_at(gx, adfv, pf, mf, b0, k + 1, nr, gvl1, ae1, g1, pv1);
        if (pv1) gm += gmin_x(g1, mf[i], pf[i + 1]);
        else gp += gmin_x(g1, pf[i], mf[i + 1]);
    }
    g_pf[i] = gp;
    g_mf[i] = gm;
}

// ---- limited horizontal flux (tracer_adv.py:460-466): factors' cotangents --------
__global__ void __launch_bounds__(256)
ka_hlim(DevMesh m, const double* __restrict__ ghsc, const double* __restrict__ adfh,
        const double* __restrict__ pf, const double* __restrict__ mf,
        double* __restrict__ g_pf, double* __restrict__ g_mf) {
    FLAT_NL(m);
    const int L = m.L;
    double ap = 0.0, am = 0.0;
    for (int p = m.ecsr_off[n], pe = m.ecsr_off[n + 1]; p < pe; ++p) {
        const int ent = m.ecsr[p];
        const int e = ent >> 1;
        const int2 nd = m.edges[e];
        const size_t i1 = (size_t)nd.x * L + k, i2 = (size_t)nd.y * L + k;
        const double adf = adfh[(size_t)e * L + k];
        const bool pos = adf >= 0.0;
        const double g_ae = (ghsc[i1] - ghsc[i2]) * adf;
        const double a1 = pos ? pf[i1] : mf[i1], a2 = pos ? mf[i2] : pf[i2];
        if ((ent & 1) == 0) {
            const double ga = gmin_x(g_ae, a1, a2);
            if (pos) ap += ga; else am += ga;
        } else {
            const double ga = gmin_x(g_ae, a2, a1);
            if (pos) am += ga; else ap += ga;
        }
    }
    g_pf[i] += ap;
    g_mf[i] += am;
}

// cluster bounds of node n at level k (zalesak_limit a4, tracer_adv.py:407-424)
__device__ __forceinline__ void cluster(const DevMesh& m, const double* tvmax,
                                        const double* tvmin, size_t i, int k, int2 nr,
                                        double& cmax, double& cmin) {
    const int L = m.L;
    const size_t im = k > 0 ? i - 1 : i;
    if ((k == nr.x) || (k == nr.y - 1)) {
        cmax = tvmax[i];
        cmin = tvmin[i];
    } else {
        cmax = fmax(fmax(tvmax[im], tvmax[i]), (k + 1 < L) ? tvmax[i + 1] : 0.0);
        cmin = fmin(fmin(tvmin[im], tvmin[i]), (k + 1 < L) ? tvmin[i + 1] : 0.0);
    }
}

// ---- limiter factors (zalesak_limit b2, tracer_adv.py:441-451) ------------------
__global__ void __launch_bounds__(256)
ka_fac(DevMesh m, const double* __restrict__ LO, const double* __restrict__ tvmax,
       const double* __restrict__ tvmin, const double* __restrict__ fplus,
       const double* __restrict__ fminus, const double* __restrict__ hnn, double dt,
       const double* __restrict__ g_pf, const double* __restrict__ g_mf,
       double* __restrict__ g_fp, double* __restrict__ g_fm, double* __restrict__ g_fmax,
       double* __restrict__ g_fmin, double* __restrict__ gLO) {
    FLAT_NL(m);
    const int2 nr = m.nlev[n];
    const bool nlm = in_layer(nr, k);
    double cmax, cmin;
    cluster(m, tvmax, tvmin, i, k, nr, cmax, cmin);
    const double lo = LO[i];
    const double ttf_max = nlm ? cmax - lo : 0.0;
    const double ttf_min = nlm ? cmin - lo : 0.0;
    const double a = m.areasvol[i], hn = hnn[i];
    const double safe_a = a > 0.0 ? a : 1.0, safe_hn = hn > 0.0 ? hn : 1.0;
    const double flux_pos = fplus[i] * dt / safe_a / safe_hn + FCT_FLUX_EPS;
    const double flux_neg = fminus[i] * dt / safe_a / safe_hn - FCT_FLUX_EPS;
    const bool valid = nlm && a > 0.0 && hn > 0.0;
    const double rp = ttf_max / flux_pos, rm = ttf_min / flux_neg;
    const double g_rp = gmin_x(valid ? g_pf[i] : 0.0, rp, 1.0);
    const double g_rm = gmin_x(valid ? g_mf[i] : 0.0, rm, 1.0);
    const double gfx = g_rp / flux_pos, gfn = g_rm / flux_neg;
    const double g_flp = (-g_rp * ttf_max) * (1.0 / (flux_pos * flux_pos));
    const double g_fln = (-g_rm * ttf_min) * (1.0 / (flux_neg * flux_neg));
    g_fp[i] = g_flp / safe_hn / safe_a * dt;
    g_fm[i] = g_fln / safe_hn / safe_a * dt;
    const double mx = nlm ? gfx : 0.0, mn = nlm ? gfn : 0.0;
    g_fmax[i] = mx;
    g_fmin[i] = mn;
    gLO[i] = gLO[i] - (mx + mn);
}

// ---- adf_h cotangent: limited-flux part and the P+/P- sums (:426-435, :460-466) --
__global__ void __launch_bounds__(256)
ka_adfh(DevMesh m, const double* __restrict__ ghsc, const double* __restrict__ adfh,
        const double* __restrict__ pf, const double* __restrict__ mf,
        const double* __restrict__ g_fp, const double* __restrict__ g_fm,
        double* __restrict__ g_adfh) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.D * m.L) return;
    const int L = m.L;
    const int e = (int)(t / L), k = (int)(t - (int64_t)e * L);
    const int2 nd = m.edges[e];
    const size_t i1 = (size_t)nd.x * L + k, i2 = (size_t)nd.y * L + k;
    const double adf = adfh[t];
    const bool pos = adf >= 0.0;
    const double ae = fmin(pos ? pf[i1] : mf[i1], pos ? mf[i2] : pf[i2]);
    double g = (ghsc[i1] - ghsc[i2]) * ae;
    const double ph0 = gmax_x(g_fp[i1], adf, 0.0), ph1 = gmax_x(g_fp[i2], -adf, 0.0);
    const double nh0 = gmin_x(g_fm[i1], adf, 0.0), nh1 = gmin_x(g_fm[i2], -adf, 0.0);
    g = g + (((ph0 - ph1) + nh0) - nh1);
    g_adfh[t] = g;
}

// ---- P+/P- vertical terms (tracer_adv.py:426-435) and the cluster bounds (a4) ------
// cotangents of a 3-level cluster at level j: (a, b, c) = (tv[j-1], tv[j], tv[j+1]),
// tv[-1] := tv[0], tv[L] := 0, combined as max(max(a, b), c) (min for the lower bound)
template <bool MAX>
__device__ __forceinline__ void cluster_vjp(const double* tv, size_t b0, int j, int L, double g,
                                            double& ga, double& gb, double& gc) {
    const double a = tv[b0 + (j > 0 ? j - 1 : 0)], b = tv[b0 + j];
    const double c = (j + 1 < L) ? tv[b0 + j + 1] : 0.0;
    if (MAX) {
        const double m1 = fmax(a, b);
        const double gm1 = gmax_x(g, m1, c);
        gc = gmax_x(g, c, m1);
        ga = gmax_x(gm1, a, b);
        gb = gmax_x(gm1, b, a);
    } else {
        const double m1 = fmin(a, b);
        const double gm1 = gmin_x(g, m1, c);
        gc = gmin_x(g, c, m1);
        ga = gmin_x(gm1, a, b);
        gb = gmin_x(gm1, b, a);
    }
}

template <bool MAX>
__device__ __forceinline__ double cluster_gather(const double* tv, const double* g, size_t b0,
                                                 int k, int L, int2 nr) {
    auto own = [&](int j) { return (j == nr.x) || (j == nr.y - 1); };
    double ga, gb, gc, out;
    if (own(k)) {
        out = g[b0 + k];
    } else {
        cluster_vjp<MAX>(tv, b0, k, L, g[b0 + k], ga, gb, gc);
        out = gb + (k == 0 ? ga : 0.0);
    }
    if (k + 1 < L && !own(k + 1)) {
        cluster_vjp<MAX>(tv, b0, k + 1, L, g[b0 + k + 1], ga, gb, gc);
        out += ga;
    }
    if (k > 0 && !own(k - 1)) {
        cluster_vjp<MAX>(tv, b0, k - 1, L, g[b0 + k - 1], ga, gb, gc);
        out += gc;
    }
    return out;
}

__global__ void __launch_bounds__(256)
ka_adfv(DevMesh m, const double* __restrict__ adfv, const double* __restrict__ g_fp,
        const double* __restrict__ g_fm, const double* __restrict__ g_fmax,
        const double* __restrict__ g_fmin, const double* __restrict__ tvmax,
        const double* __restrict__ tvmin, double* __restrict__ g_adfv,
        double* __restrict__ g_fv, double* __restrict__ g_tvx, double* __restrict__ g_tvn) {
    FLAT_NL(m);
    const int L = m.L;
    const int2 nr = m.nlev[n];
    const size_t b0 = (size_t)n * L;
    const bool nlm = in_layer(nr, k);
    const double f = adfv[i];
    double g = g_adfv[i];
    g += gmax_x(nlm ? g_fp[i] : 0.0, f, 0.0) + gmin_x(nlm ? g_fm[i] : 0.0, f, 0.0);
    if (k > 0) {   // max(-adf_v[k], 0) in P+ of level k-1
        const bool nlm1 = in_layer(nr, k - 1);
        const double gpv = nlm1 ? g_fp[i - 1] : 0.0, gnv = nlm1 ? g_fm[i - 1] : 0.0;
        g += -gmax_x(gpv, -f, 0.0) - gmin_x(gnv, -f, 0.0);
    }
    g_adfv[i] = g;
    g_fv[i] = -g;
    g_tvx[i] = cluster_gather<true>(tvmax, g_fmax, b0, k, L, nr);
    g_tvn[i] = cluster_gather<false>(tvmin, g_fmin, b0, k, L, nr);
}

// ---- QR4C backward (adv_flux_ver_ho, tracer_adv.py:241-295) --------------------------
// contributions of interface j to A[j], A[j-1], A[j-2], A[j+1] and to w[j]
struct QrTerms {
    double a0 = 0.0, am1 = 0.0, am2 = 0.0, ap1 = 0.0, w = 0.0;
};

// QR4C inputs that do not depend on the tracer: the vertical velocity of adf_v (w_e, or
// w_e + w_i under the implicit split) and the moving-level stencil (Z3d, zbar3), or
// NULL for the mesh's static one
struct QrFlow {
    const double *w, *w_i, *Z3d, *zbar3;
};

__device__ __forceinline__ QrTerms qr_vjp(const DevMesh& m, const double* T, const double* Told,
                                          const QrFlow& f, const double* g_adfv, size_t b0,
                                          int j, int2 nr) {
    QrTerms r;
    const int L = m.L;
    if (!in_iface(nr, j) || j == nr.y) return r;
    const size_t i = b0 + j, im = b0 + (j > 0 ? j - 1 : 0);
    const double gav = g_adfv[i];
    const double W = f.w_i ? f.w[i] + f.w_i[i] : f.w[i], ar = m.area[i];
    const double A = ab2a(Told[i], T[i]);
    const double Am1 = ab2a(Told[im], T[im]);
    if (j == nr.x) {
        r.a0 = -(gav * ar * W);
        r.w = gav * ar * (-A);
    } else if (j == nr.x + 1 || j == nr.y - 1) {
        const double gcm = gav * ar;
        const double gs = -0.5 * (gcm * W);
        r.w = gcm * (-0.5 * (Am1 + A));
        r.a0 = gs;
        r.am1 = gs;
    } else if (j >= nr.x + 2 && j <= nr.y - 2) {
        const size_t im2 = b0 + (j > 1 ? j - 2 : 0);
        const size_t ip = b0 + (j + 1 < L ? j + 1 : L - 1);
        const double Am2 = ab2a(Told[im2], T[im2]);
        const double Ap1 = (j + 1 < L) ? ab2a(Told[ip], T[ip]) : 0.0;
        double Znz, Zm1, Zm2, Zp1, zb;
        if (f.Z3d) {
            Znz = f.Z3d[i]; Zm1 = f.Z3d[im]; Zm2 = f.Z3d[im2]; Zp1 = f.Z3d[ip];
            zb = f.zbar3[i];
        } else {
            Znz = m.Zs[j]; Zm1 = m.Zs[L + j]; Zm2 = m.Zs[2 * L + j]; Zp1 = m.Zs[3 * L + j];
            zb = m.zbar[j];
        }
        const double dc = nz_safe(Zm1 - Znz), du = nz_safe(Znz - Zp1), dd = nz_safe(Zm2 - Zm1);
        const double qc = (Am1 - A) / dc, qu = (A - Ap1) / du, qd = (Am2 - Am1) / dd;
        const double Tmean1 = A + (2.0 * qc + qu) * (zb - Znz) / 3.0;
        const double Tmean2 = Am1 + (2.0 * qc + qd) * (zb - Zm1) / 3.0;
        const double gqm = gav * ar;
        r.w = gqm * (-0.5 * (Tmean1 + Tmean2));
        const double gtm = -0.5 * (gqm * W);
        const double gr1 = gtm / 3.0 * (zb - Znz), gr2 = gtm / 3.0 * (zb - Zm1);
        const double gqc = 2.0 * gr1 + 2.0 * gr2;
        double gAk = gtm, gAm1 = gtm;
        gAm1 += gqc / dc;
        gAk -= gqc / dc;
        gAk += gr1 / du;
        r.ap1 = (j + 1 < L) ? -(gr1 / du) : 0.0;
        r.am2 = gr2 / dd;
        gAm1 -= gr2 / dd;
        r.a0 = gAk;
        r.am1 = gAm1;
    }
    return r;
}

// A[k] gathers the terms of interfaces k, k+1, k+2 and k-1; w[k] those of interface k
// (to w_e and, under the split, to w_i as well)
__global__ void __launch_bounds__(256)
ka_qr(DevMesh m, const double* __restrict__ T, const double* __restrict__ Told, QrFlow f,
      const double* __restrict__ g_adfv, double* __restrict__ gA, double* __restrict__ gw,
      double* __restrict__ gwi) {
    FLAT_NL(m);
    const int L = m.L;
    const int2 nr = m.nlev[n];
    const size_t b0 = (size_t)n * L;
    const QrTerms t0 = qr_vjp(m, T, Told, f, g_adfv, b0, k, nr);
    double ga = t0.a0;
    if (k + 1 < L) ga += qr_vjp(m, T, Told, f, g_adfv, b0, k + 1, nr).am1;
    if (k + 2 < L) ga += qr_vjp(m, T, Told, f, g_adfv, b0, k + 2, nr).am2;
    if (k > 0) ga += qr_vjp(m, T, Told, f, g_adfv, b0, k - 1, nr).ap1;
    gA[i] = ga;
    gw[i] += t0.w;
    if (gwi) gwi[i] += t0.w;
}

// element extremum of max/min(LO, T) over its vertices at level k (a2)
__device__ __forceinline__ void elem_aux(const DevMesh& m, const double* LO, const double* T,
                                         int e, int k, double& amax, double& amin) {
    const int L = m.L;
    amax = -FCT_BIG;
    amin = FCT_BIG;
    if (!in_layer(m.elev[e], k)) return;
    const size_t j0 = (size_t)m.en[3 * e] * L + k, j1 = (size_t)m.en[3 * e + 1] * L + k,
                 j2 = (size_t)m.en[3 * e + 2] * L + k;
    amax = fmax(fmax(fmax(LO[j0], T[j0]), fmax(LO[j1], T[j1])), fmax(LO[j2], T[j2]));
    amin = fmin(fmin(fmin(LO[j0], T[j0]), fmin(LO[j1], T[j1])), fmin(LO[j2], T[j2]));
}

// ---- a3 segment extremum: the node's cotangent shared by its tied elements --------
__global__ void __launch_bounds__(256)
ka_share(DevMesh m, const double* __restrict__ LO, const double* __restrict__ T,
         const double* __restrict__ tvmax, const double* __restrict__ tvmin,
         const double* __restrict__ g_tvx, const double* __restrict__ g_tvn,
         double* __restrict__ sh_x, double* __restrict__ sh_n) {
    FLAT_NL(m);
    int cx = 0, cn = 0;
    for (int p = m.ncsr_off[n], pe = m.ncsr_off[n + 1]; p < pe; ++p) {
        const int e = m.ncsr[p];
        // an element appears once per vertex it shares with n (once in a valid mesh)
        double amax, amin;
        elem_aux(m, LO, T, e, k, amax, amin);
        cx += amax == tvmax[i];
        cn += amin == tvmin[i];
    }
    sh_x[i] = cx > 0 ? g_tvx[i] / (double)cx : 0.0;
    sh_n[i] = cn > 0 ? g_tvn[i] / (double)cn : 0.0;
}

// ---- a2 + a1: vertex cotangents of max/min(LO, T) ---------------------------------
__global__ void __launch_bounds__(256)
ka_vertex(DevMesh m, const double* __restrict__ LO, const double* __restrict__ T,
          const double* __restrict__ tvmax, const double* __restrict__ tvmin,
          const double* __restrict__ sh_x, const double* __restrict__ sh_n,
          double* __restrict__ gLO, double* __restrict__ gT) {
    FLAT_NL(m);
    const int L = m.L;
    double gmx = 0.0, gmn = 0.0;
    const double wx_v = fmax(LO[i], T[i]), wn_v = fmin(LO[i], T[i]);
    for (int p = m.ncsr_off[n], pe = m.ncsr_off[n + 1]; p < pe; ++p) {
        const int e = m.ncsr[p];
        if (!in_layer(m.elev[e], k)) continue;
        double amax, amin;
        elem_aux(m, LO, T, e, k, amax, amin);
        double gax = 0.0, gan = 0.0;
        int hx = 0, hn = 0;
        for (int v = 0; v < 3; ++v) {
            const size_t j = (size_t)m.en[3 * e + v] * L + k;
            gax = gax + (amax == tvmax[j] ? sh_x[j] : 0.0);
            gan = gan + (amin == tvmin[j] ? sh_n[j] : 0.0);
            hx += fmax(LO[j], T[j]) == amax;
            hn += fmin(LO[j], T[j]) == amin;
        }
        if (wx_v == amax) gmx = gmx + gax / (double)hx;
        if (wn_v == amin) gmn = gmn + gan / (double)hn;
    }
    const double lo = LO[i], t = T[i];
    gLO[i] = (gLO[i] + gmax_x(gmx, lo, t)) + gmin_x(gmn, lo, t);
    gT[i] = (gT[i] + gmax_x(gmx, t, lo)) + gmin_x(gmn, t, lo);
}

// ---- LO backward (tracer_adv.py:298-313) ---------------------------------------------
__device__ __forceinline__ double lo_div(const DevMesh& m, const double* gLO, const double* hnn,
                                         double dt, size_t i, int k, int2 nr) {
    const double a = m.areasvol[i], hn = hnn[i];
    const double safe_a = a > 0.0 ? a : 1.0, safe_hn = hn > 0.0 ? hn : 1.0;
    const bool valid = in_layer(nr, k) && hn > 0.0 && a > 0.0;
    return (valid ? gLO[i] : 0.0) / safe_hn / safe_a * dt;
}

__global__ void __launch_bounds__(256)
ka_lo(DevMesh m, const double* __restrict__ hnode, const double* __restrict__ hnn, double dt,
      const double* __restrict__ gLO, double* __restrict__ g_div, double* __restrict__ gT) {
    FLAT_NL(m);
    const int2 nr = m.nlev[n];
    const double a = m.areasvol[i], hn = hnn[i];
    const double safe_hn = hn > 0.0 ? hn : 1.0;
    const bool valid = in_layer(nr, k) && hn > 0.0 && a > 0.0;
    const double gnum = (valid ? gLO[i] : 0.0) / safe_hn;
    gT[i] += gnum * hnode[i];
    g_div[i] = lo_div(m, gLO, hnn, dt, i, k, nr);
}

// ---- LO vertical flux (adv_flux_ver, tracer_adv.py:108-120) --------------------------
// fv[k] = mask(-0.5 (T[k] (w+|w|) + T[k-1] (w-|w|)) * area), T[-1] := T[0]; its
// cotangent is -g_adf_v + (g_div[k] - g_div[k-1]) (the LO divergence fv[k] - fv[k+1])
__device__ __forceinline__ double fv_cot(const DevMesh& m, const double* g_fv, const double* g_div,
                                         size_t b0, int k, int2 nr) {
    const size_t i = b0 + k;
    const double gd = g_div[i];
    const double g = g_fv[i] + (k > 0 ? gd + (-g_div[i - 1]) : gd);
    return -0.5 * ((in_layer(nr, k) ? g : 0.0) * m.area[i]);
}

__global__ void __launch_bounds__(256)
ka_fv(DevMesh m, const double* __restrict__ T, const double* __restrict__ w,
      const double* __restrict__ g_fv, const double* __restrict__ g_div,
      double* __restrict__ gT, double* __restrict__ gw) {
    FLAT_NL(m);
    const int L = m.L;
    const int2 nr = m.nlev[n];
    const size_t b0 = (size_t)n * L, im = b0 + (k > 0 ? k - 1 : 0);
    const double gs = fv_cot(m, g_fv, g_div, b0, k, nr);
    const double W = w[i], aw = fabs(W);
    double t = gs * (W + aw);
    if (k == 0) t += gs * (W - aw);
    if (k + 1 < L) {
        const double W1 = w[i + 1];
        t += fv_cot(m, g_fv, g_div, b0, k + 1, nr) * (W1 - fabs(W1));
    }
    gT[i] += t;
    const double gwp = gs * T[i], gwm = gs * T[im];
    gw[i] += gwp + gwm + gabs(gwp - gwm, W);
}

// The same under the implicit split, where the two fluxes differ: LO's uses w_e and
// takes the LO-divergence cotangent, adf_v's uses w_e + w_i and takes -g_adf_v
// (tracer_adv.py:551-552, 568-572).
__device__ __forceinline__ double fv_cot_part(const DevMesh& m, const double* g, size_t b0, int k,
                                              int2 nr, bool div) {
    const size_t i = b0 + k;
    const double gd = g[i];
    const double v = div ? (k > 0 ? gd + (-g[i - 1]) : gd) : gd;
    return -0.5 * ((in_layer(nr, k) ? v : 0.0) * m.area[i]);
}

__global__ void __launch_bounds__(256)
ka_fv_split(DevMesh m, const double* __restrict__ T, const double* __restrict__ w_e,
            const double* __restrict__ w_i, const double* __restrict__ g_fv,
            const double* __restrict__ g_div, double* __restrict__ gT, double* __restrict__ gw,
            double* __restrict__ gwi) {
    FLAT_NL(m);
    const int L = m.L;
    const int2 nr = m.nlev[n];
    const size_t b0 = (size_t)n * L, im = b0 + (k > 0 ? k - 1 : 0);
    const double gl = fv_cot_part(m, g_div, b0, k, nr, true);
    const double gh = fv_cot_part(m, g_fv, b0, k, nr, false);
    const double We = w_e[i], Wf = w_e[i] + w_i[i];
    const double ae = fabs(We), af = fabs(Wf);
    double t = gl * (We + ae) + gh * (Wf + af);
    if (k == 0) t += gl * (We - ae) + gh * (Wf - af);
    if (k + 1 < L) {
        const double We1 = w_e[i + 1], Wf1 = w_e[i + 1] + w_i[i + 1];
        t += fv_cot_part(m, g_div, b0, k + 1, nr, true) * (We1 - fabs(We1)) +
             fv_cot_part(m, g_fv, b0, k + 1, nr, false) * (Wf1 - fabs(Wf1));
    }
    gT[i] += t;
    const double lp = gl * T[i], lm = gl * T[im];
    const double hp = gh * T[i], hm = gh * T[im];
    const double gf = hp + hm + gabs(hp - hm, Wf);
    gw[i] += (lp + lm + gabs(lp - lm, We)) + gf;
    gwi[i] += gf;
}

// ---- adv_tra_vert_impl on LO (tracer_adv.py:555-556, 487-529) ------------------------
// One thread per column: the Thomas recurrence of ops.tdma (ops.py:121-168) is rerun
// from the tape's LO before the solve (X), then reversed step by step as jax.grad
// reverses its two scans. On entry gX holds the cotangent of the solved LO; on exit the
// cotangent of X. The w_i cotangent is added to gwi. cp, dp, xs, gd: [N*L] scratch.
__global__ void __launch_bounds__(256)
ka_vimpl(DevMesh m, const double* __restrict__ w_i, const double* __restrict__ hnn, double dt,
         const double* __restrict__ X, double* __restrict__ gX, double* __restrict__ gwi,
         double* __restrict__ cp, double* __restrict__ dp, double* __restrict__ xs,
         double* __restrict__ gd) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= m.N) return;
    const int L = m.L;
    const int2 nr = m.nlev[n];
    const size_t b0 = (size_t)n * L;
    auto rhs_at = [&](const VimplRow& r, int k) {
        const size_t i = b0 + k;
        const double Td = X[k > 0 ? i - 1 : i], Tc = X[i], Tu = k + 1 < L ? X[i + 1] : 0.0;
        return r.valid ? -(r.a * Td + r.b_adv * Tc + r.c * Tu) : 0.0;
    };
    // forward sweep and back substitution (x kept in xs)
    double cprev = 0.0, dprev = 0.0;
    for (int k = 0; k < L; ++k) {
        const VimplRow r = vimpl_row(m, w_i, hnn, dt, b0, k, nr);
        const double mm = r.b - r.a * cprev;
        cprev = r.c / mm;
        dprev = (rhs_at(r, k) - r.a * dprev) / mm;
        cp[b0 + k] = cprev;
        dp[b0 + k] = dprev;
    }
    double xn = 0.0;
    for (int k = L - 1; k >= 0; --k) {
        xn = dp[b0 + k] - cp[b0 + k] * xn;
        xs[b0 + k] = xn;
    }
    // back substitution reversed: x[k] = dp[k] - cp[k] x[k+1]; gd <- cot of dp,
    // xs <- cot of cp (x[k+1] is read before level k+1 is overwritten)
    double carry = 0.0;
    for (int k = 0; k < L; ++k) {
        const size_t i = b0 + k;
        const double G = (in_layer(nr, k) ? gX[i] : 0.0) + carry;
        gd[i] = G;
        const double x1 = k + 1 < L ? xs[i + 1] : 0.0;
        xs[i] = -G * x1;
        carry = -G * cp[i];
    }
    // forward sweep reversed; gX[k] keeps the identity term of X + dT and gathers the
    // right-hand side's transpose; gwi[k] gathers the upper face (w_i[k]) of row k and
    // the lower face of row k-1
    double gcc = 0.0, gdc = 0.0;
    for (int k = L - 1; k >= 0; --k) {
        const size_t i = b0 + k;
        const VimplRow r = vimpl_row(m, w_i, hnn, dt, b0, k, nr);
        const double cpv = k > 0 ? cp[i - 1] : 0.0, dpv = k > 0 ? dp[i - 1] : 0.0;
        const double mm = r.b - r.a * cpv;
        const double rhs = rhs_at(r, k);
        const double num = rhs - r.a * dpv;
        const double GC = xs[i] + gcc, GD = gd[i] + gdc;
        const double g_num = GD / mm;
        double g_m = (-GD * num) * (1.0 / (mm * mm));
        double g_a = -g_num * dpv;
        gdc = -g_num * r.a;
        double g_c = GC / mm;
        g_m = g_m + (-GC * r.c) * (1.0 / (mm * mm));
        const double g_b = g_m;
        g_a += -g_m * cpv;
        gcc = -g_m * r.a;
        // rhs = where(valid, -(a X[k-1] + b_adv X[k] + c X[k+1]), 0)
        const double gs = -(r.valid ? g_num : 0.0);
        const double Td = X[k > 0 ? i - 1 : i], Tc = X[i], Tu = k + 1 < L ? X[i + 1] : 0.0;
        g_a += gs * Td;
        double g_badv = gs * Tc;
        g_c += gs * Tu;
        if (k > 0) gX[i - 1] += gs * r.a; else gX[i] += gs * r.a;
        gX[i] += gs * r.b_adv;
        if (k + 1 < L) gX[i + 1] += gs * r.c;
        g_badv += r.valid ? g_b : 0.0;
        const double g_aadv = (r.valid ? g_a : 0.0) * r.v_top;
        const double g_cadv = (r.valid ? g_c : 0.0) * r.v_bot;
        double g_wu = r.surf ? 0.0 : gmin_x(g_aadv, r.wu, 0.0);
        const double gt = g_badv * r.v_top;
        g_wu += r.surf ? gt : gmax_x(gt, r.wu, 0.0);
        const double gbb = g_badv * r.v_bot;
        double g_wd = r.bot ? 0.0 : -gmin_x(gbb, r.wd, 0.0);
        g_wd += r.bot ? 0.0 : -gmax_x(g_cadv, r.wd, 0.0);
        // wd of row k is w_i[k+1] (0 on the last level)
        if (k + 1 < L) gwi[i + 1] += g_wd;
        gwi[i] += g_wu;
    }
}

// ---- edges: adf_h = HO(A, eud) - fh(T) and fh's LO-divergence cotangent ------------
__global__ void __launch_bounds__(256)
ka_edge(DevMesh m, const double* __restrict__ T, const double* __restrict__ Told,
        const double* __restrict__ vflux, const double2* __restrict__ tr,
        const double2* __restrict__ gnode, const double* __restrict__ g_adfh,
        const double* __restrict__ g_div, double* __restrict__ g_vf, double* __restrict__ eA,
        double* __restrict__ eT, double4* __restrict__ eud) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.D * m.L) return;
    const int L = m.L;
    const int e = (int)(t / L), k = (int)(t - (int64_t)e * L);
    const int2 nd = m.edges[e];
    const size_t i1 = (size_t)nd.x * L + k, i2 = (size_t)nd.y * L + k;
    const double vf = vflux[t], av = fabs(vf);
    const double A1 = ab2a(Told[i1], T[i1]), A2 = ab2a(Told[i2], T[i2]);
    // eud: up/dn triangle gradients inside the shared range, else node averages
    const int2 ur = m.utri_rng[e];
    const bool tri = k >= ur.x && k < ur.y;
    const int2 ud = m.updn[e];
    const double2 gu = tri ? tr[(size_t)ud.x * L + k] : gnode[i1];
    const double2 gd = tri ? tr[(size_t)ud.y * L + k] : gnode[i2];
    const double2 ec = m.edge_c[e];
    const double diff = A2 - A1;
    const double Tm1 = A1 + (2.0 * diff + ec.x * gu.x + ec.y * gu.y) / 6.0;
    const double Tm2 = A2 - (2.0 * diff + ec.x * gd.x + ec.y * gd.y) / 6.0;
    const double gh = g_adfh[t];
    const double gs2 = -0.5 * gh;
    const double gTm1 = gs2 * (vf + av), gTm2 = gs2 * (vf - av);
    double gp = gs2 * Tm1, gm = gs2 * Tm2;
    double gv = gp + gm + gabs(gp - gm, vf);
    const double gi1 = gTm1 / 6.0, gi2 = -(gTm2 / 6.0);
    eA[2 * t] = gTm1 - 2.0 * (gi1 + gi2);
    eA[2 * t + 1] = gTm2 + 2.0 * (gi1 + gi2);
    eud[t] = make_double4(gi1 * ec.x, gi2 * ec.x, gi1 * ec.y, gi2 * ec.y);
    // fh = -0.5 (T1 (v+|v|) + T2 (v-|v|)): from -adf_h and the LO divergence
    const double gfh = -gh + (g_div[i1] - g_div[i2]);
    const double gs1 = -0.5 * gfh;
    eT[2 * t] = gs1 * (vf + av);
    eT[2 * t + 1] = gs1 * (vf - av);
    gp = gs1 * T[i1];
    gm = gs1 * T[i2];
    gv = gv + (gp + gm + gabs(gp - gm, vf));
    g_vf[t] += gv;
}

// ---- nodes gather their edges' contributions to A, T and the node gradients -------
__global__ void __launch_bounds__(256)
ka_node_edges(DevMesh m, const double* __restrict__ eA, const double* __restrict__ eT,
              const double4* __restrict__ eud, double* __restrict__ gA, double* __restrict__ gT,
              double2* __restrict__ g_num) {
    FLAT_NL(m);
    const int L = m.L;
    double ga = 0.0, gt = 0.0, gx = 0.0, gy = 0.0;
    for (int p = m.ecsr_off[n], pe = m.ecsr_off[n + 1]; p < pe; ++p) {
        const int ent = m.ecsr[p];
        const int e = ent >> 1, side = ent & 1;
        const size_t t = (size_t)e * L + k;
        ga = ga + eA[2 * t + side];
        gt = gt + eT[2 * t + side];
        const int2 ur = m.utri_rng[e];
        if (!(k >= ur.x && k < ur.y)) {
            const double4 u = eud[t];
            gx = gx + (side ? u.y : u.x);
            gy = gy + (side ? u.w : u.z);
        }
    }
    gA[i] += ga;
    gT[i] += gt;
    // the node average's denominator (tracer_adv.py:333-341)
    double den = 0.0;
    for (int p = m.ncsr_off[n], pe = m.ncsr_off[n + 1]; p < pe; ++p) {
        const int e = m.ncsr[p];
        den = den + (in_layer(m.elev[e], k) ? m.earea[e] : 0.0);
    }
    const double d = den > 0.0 ? den : 1.0;
    g_num[i] = make_double2(gx / d, gy / d);
}

// ---- element gradient cotangents (tracer_adv.py:316-368) ---------------------------
__global__ void __launch_bounds__(256)
ka_tr(DevMesh m, const int* __restrict__ ud_off, const int* __restrict__ ud_ent,
      const double4* __restrict__ eud, const double2* __restrict__ g_num,
      double2* __restrict__ g_tr) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.E * m.L) return;
    const int L = m.L;
    const int e = (int)(t / L), k = (int)(t - (int64_t)e * L);
    double gx = 0.0, gy = 0.0;
    // edges using this triangle as their up (role 0) or down (role 1) triangle
    for (int p = ud_off[e], pe = ud_off[e + 1]; p < pe; ++p) {
        const int ent = ud_ent[p];
        const int d = ent >> 1, role = ent & 1;
        const int2 ur = m.utri_rng[d];
        if (!(k >= ur.x && k < ur.y)) continue;
        const double4 u = eud[(size_t)d * L + k];
        gx = gx + (role ? u.y : u.x);
        gy = gy + (role ? u.w : u.z);
    }
    const bool em = in_layer(m.elev[e], k);
    const double aw = em ? m.earea[e] : 0.0;
    double sx = 0.0, sy = 0.0;
    for (int v = 0; v < 3; ++v) {
        const double2 g = g_num[(size_t)m.en[3 * e + v] * L + k];
        sx = sx + g.x;
        sy = sy + g.y;
    }
    gx = gx + sx * aw;
    gy = gy + sy * aw;
    g_tr[t] = em ? make_double2(gx, gy) : make_double2(0.0, 0.0);
}

// ---- nodes gather the gradient cotangents; AB2 (tracer_adv.py:59-66) --------------
__global__ void __launch_bounds__(256)
ka_node_final(DevMesh m, const double2* __restrict__ g_tr, const double* __restrict__ gA,
              double* __restrict__ gT, double* __restrict__ gTold) {
    FLAT_NL(m);
    const int L = m.L;
    double acc = gT[i];
    for (int p = m.ncsr_off[n], pe = m.ncsr_off[n + 1]; p < pe; ++p) {
        const int e = m.ncsr[p];
        const double2 g = g_tr[(size_t)e * L + k];
        for (int v = 0; v < 3; ++v) {
            if (m.en[3 * e + v] != n) continue;
            acc = acc + (g.x * m.gs[6 * (size_t)e + v] + g.y * m.gs[6 * (size_t)e + 3 + v]);
        }
    }
    const double ga = gA[i];
    gT[i] = acc + AB_NEW * ga;
    gTold[i] = AB_OLD * ga;
}

// ---- uv cotangent from the vflux cotangent (edge_vflux, tracer_adv.py:69-93) ------
__global__ void __launch_bounds__(256)
ka_uv(DevMesh m, const int* __restrict__ et_off, const int* __restrict__ et_ent,
      const double* __restrict__ helem, const double* __restrict__ g_vf,
      double2* __restrict__ g_uv) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.E * m.L) return;
    const int L = m.L;
    const int e = (int)(t / L), k = (int)(t - (int64_t)e * L);
    double gx = 0.0, gy = 0.0;
    if (in_layer(m.elev[e], k)) {
        const double h = helem[t];
        for (int p = et_off[e], pe = et_off[e + 1]; p < pe; ++p) {
            const int ent = et_ent[p];
            const int d = ent >> 1;
            const double4 c = m.cross[d];
            const double g = g_vf[(size_t)d * L + k] * h;
            if ((ent & 1) == 0) {
                gx = gx + g * c.y;
                gy = gy + (-(g * c.x));
            } else {
                gy = gy + g * c.z;
                gx = gx + (-(g * c.w));
            }
        }
    }
    g_uv[t] = make_double2(gx, gy);
}

#undef FLAT_NL

__global__ void __launch_bounds__(256) ka_zero(double* __restrict__ p, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0.0;
}

inline unsigned grid(int64_t count) { return (unsigned)((count + 255) / 256); }

}  // namespace

int launch_backward(const DevMesh& m, const AdjointWork& w, const fct_step_args* a,
                    const double* const* g_out, double* const* gT, double* const* gTold,
                    double* g_uv, double* g_w, double* g_wi, hipStream_t s) {
    const int64_t NL = (int64_t)m.N * m.L, DL = (int64_t)m.D * m.L, EL = (int64_t)m.E * m.L;
    const dim3 b(256);
    const bool split = a->w_i != nullptr;
    // zero fills by kernel (no memset nodes if a caller captures this in a graph)
    double* gw = g_w ? g_w : w.gw;
    double* gwi = split ? (g_wi ? g_wi : w.gwi) : nullptr;
    hipLaunchKernelGGL(ka_zero, dim3(grid(DL)), b, 0, s, w.g_vf, DL);
    hipLaunchKernelGGL(ka_zero, dim3(grid(NL)), b, 0, s, gw, NL);
    if (split) hipLaunchKernelGGL(ka_zero, dim3(grid(NL)), b, 0, s, gwi, NL);
    int rc = launch_vflux(m, a, w.vflux, s);
    if (rc) return rc;
    fct_step_args explicit_part = *a;   // LO before the implicit solve
    explicit_part.w_i = nullptr;
    const QrFlow qf{a->w_e, a->w_i, a->Z3d, a->zbar3};
    for (int j = 0; j < a->ntracers; ++j) {
        // forward tape of tracer j (phased kernels: dense intermediates); under the split
        // LO before the solve is kept in LOx and solved into LO (sh_x, sh_n as scratch)
        if (split) {
            if ((rc = launch_lo(m, &explicit_part, j, w.vflux, w.LOx, nullptr, nullptr, s)))
                return rc;
            if (hipMemcpyAsync(w.LO, w.LOx, NL * sizeof(double), hipMemcpyDeviceToDevice, s) !=
                hipSuccess)
                return set_error(FCT_E_HIP, "adjoint tape copy failed");
            if ((rc = launch_vert_impl(m, a->w_i, a->hnode_new, a->dt, w.LO, w.sh_x, w.sh_n, s)))
                return rc;
        } else if ((rc = launch_lo(m, a, j, w.vflux, w.LO, nullptr, nullptr, s))) {
            return rc;
        }
        if ((rc = launch_grad(m, a, j, w.trxy, s)) ||
            (rc = launch_pm(m, a, j, w.vflux, w.LO, w.trxy, w.gnode, w.adfh, w.adfv, w.tvmax,
                            w.tvmin, w.fplus, w.fminus, s)))
            return rc;
        if ((rc = launch_fac(m, a, w.LO, w.tvmax, w.tvmin, w.fplus, w.fminus, w.pf, w.mf, s)))
            return rc;
        // backward
        hipLaunchKernelGGL(ka_final, dim3(grid(NL)), b, 0, s, m, g_out[j], a->hnode,
                           a->hnode_new, a->dt, gT[j], w.gLO, w.gx, w.ghsc);
        hipLaunchKernelGGL(ka_vlim, dim3(grid(NL)), b, 0, s, m, w.gx, w.adfv, w.pf, w.mf,
                           w.g_adfv, w.g_pf, w.g_mf);
        hipLaunchKernelGGL(ka_hlim, dim3(grid(NL)), b, 0, s, m, w.ghsc, w.adfh, w.pf, w.mf,
                           w.g_pf, w.g_mf);
        hipLaunchKernelGGL(ka_fac, dim3(grid(NL)), b, 0, s, m, w.LO, w.tvmax, w.tvmin, w.fplus,
                           w.fminus, a->hnode_new, a->dt, w.g_pf, w.g_mf, w.g_fp, w.g_fm,
                           w.g_fmax, w.g_fmin, w.gLO);
        hipLaunchKernelGGL(ka_adfh, dim3(grid(DL)), b, 0, s, m, w.ghsc, w.adfh, w.pf, w.mf,
                           w.g_fp, w.g_fm, w.g_adfh);
        hipLaunchKernelGGL(ka_adfv, dim3(grid(NL)), b, 0, s, m, w.adfv, w.g_fp, w.g_fm,
                           w.g_fmax, w.g_fmin, w.tvmax, w.tvmin, w.g_adfv, w.g_fv, w.g_tvx,
                           w.g_tvn);
        hipLaunchKernelGGL(ka_qr, dim3(grid(NL)), b, 0, s, m, a->T[j], a->T_old[j], qf,
                           w.g_adfv, w.gA, gw, gwi);
        hipLaunchKernelGGL(ka_share, dim3(grid(NL)), b, 0, s, m, w.LO, a->T[j], w.tvmax,
                           w.tvmin, w.g_tvx, w.g_tvn, w.sh_x, w.sh_n);
        hipLaunchKernelGGL(ka_vertex, dim3(grid(NL)), b, 0, s, m, w.LO, a->T[j], w.tvmax,
                           w.tvmin, w.sh_x, w.sh_n, w.gLO, gT[j]);
        if (split)   // the solve's transpose (sh_x, sh_n, g_tvx, g_tvn are free here)
            hipLaunchKernelGGL(ka_vimpl, dim3((unsigned)((m.N + 255) / 256)), b, 0, s, m, a->w_i,
                               a->hnode_new, a->dt, w.LOx, w.gLO, gwi, w.sh_x, w.sh_n, w.g_tvx,
                               w.g_tvn);
        hipLaunchKernelGGL(ka_lo, dim3(grid(NL)), b, 0, s, m, a->hnode, a->hnode_new, a->dt,
                           w.gLO, w.g_div, gT[j]);
        if (split)
            hipLaunchKernelGGL(ka_fv_split, dim3(grid(NL)), b, 0, s, m, a->T[j], a->w_e, a->w_i,
                               w.g_fv, w.g_div, gT[j], gw, gwi);
        else
            hipLaunchKernelGGL(ka_fv, dim3(grid(NL)), b, 0, s, m, a->T[j], a->w_e, w.g_fv,
                               w.g_div, gT[j], gw);
        hipLaunchKernelGGL(ka_edge, dim3(grid(DL)), b, 0, s, m, a->T[j], a->T_old[j], w.vflux,
                           (const double2*)w.trxy, (const double2*)w.gnode, w.g_adfh, w.g_div,
                           w.g_vf, w.eA, w.eT, (double4*)w.eud);
        hipLaunchKernelGGL(ka_node_edges, dim3(grid(NL)), b, 0, s, m, w.eA, w.eT,
                           (const double4*)w.eud, w.gA, gT[j], (double2*)w.g_num);
        hipLaunchKernelGGL(ka_tr, dim3(grid(EL)), b, 0, s, m, w.ud_off, w.ud_ent,
                           (const double4*)w.eud, (const double2*)w.g_num, (double2*)w.g_tr);
        hipLaunchKernelGGL(ka_node_final, dim3(grid(NL)), b, 0, s, m, (const double2*)w.g_tr,
                           w.gA, gT[j], gTold[j]);
        if ((rc = check_launch("adjoint kernels"))) return rc;
    }
    if (g_uv)
        hipLaunchKernelGGL(ka_uv, dim3(grid(EL)), b, 0, s, m, w.et_off, w.et_ent, a->helem,
                           w.g_vf, (double2*)g_uv);
    return check_launch("adjoint uv");
}

}  // namespace fct
