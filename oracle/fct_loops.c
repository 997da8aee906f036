/* CPU oracle, test infrastructure only: a plain-C loop restatement of one FCT tracer
 * step with the semantics of the original C kernel, fesom_tracer_advect_one_fct, as the
 * reference's own independent loop oracle restates them (fesom_jax/tests/
 * test_tracers.py:325-516, compared there with the JAX step at 1e-11 on smooth fields
 * and 1e-9 with the limiter active, :560-596).
 *
 * It is a second, independent statement of the algorithm that the numpy oracle
 * (oracle/fct_numpy.py, which follows the vectorised fesom_jax/tracer_adv.py) is checked
 * against: scalar loops over edges, nodes and elements with explicit level ranges and
 * in-place accumulation, instead of masks and segment sums. Only tests/ and bench.py's
 * CPU baseline load it; the product never does. Built by oracle/build.py with
 * -ffp-contract=off into oracle/_build/libfct_loops.so.
 *
 * Layout: row-major [entity][nl] fields, uv [elem][nl][2], 0-based indices with -1
 * sentinels, level counts 1-based as in the mesh bundle (fesom_jax/mesh.py:13-33).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int nod2D, elem2D, edge2D, nl;
    const int *edges, *edge_tri, *edge_up_dn_tri, *elem_nodes;
    const int *ulevels, *nlevels;                       /* per element */
    const int *ulevels_nod2D, *nlevels_nod2D;           /* per node */
    const int *ulevels_nod2D_max, *nlevels_nod2D_min;   /* per node */
    const int *nie_off, *nie;                           /* nod_in_elem2D CSR */
    const double *edge_cross_dxdy, *edge_dxdy, *gradient_sca;
    const double *elem_area, *elem_cos, *area, *areasvol, *Z, *zbar;
} loop_mesh;

#define R_EARTH 6367500.0
#define FLUX_EPS 1e-16
#define BIG 1e3
#define AB_OLD (-(0.5 + 0.1))
#define AB_NEW (1.5 + 0.1)

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

/* volume flux through edge e at level k: left triangle, then right (tracer_adv.py:69-83) */
static double edge_vflux(const loop_mesh* m, const double* uv, const double* helem, int e, int k) {
    const int nl = m->nl;
    const int t1 = m->edge_tri[2 * e], t2 = m->edge_tri[2 * e + 1];
    const double* cr = m->edge_cross_dxdy + 4 * (size_t)e;
    double vf = 0.0;
    if (t1 >= 0 && k >= m->ulevels[t1] - 1 && k < m->nlevels[t1] - 1) {
        const double u = uv[((size_t)t1 * nl + k) * 2], v = uv[((size_t)t1 * nl + k) * 2 + 1];
        vf += (u * cr[1] - v * cr[0]) * helem[(size_t)t1 * nl + k];
    }
    if (t2 >= 0 && k >= m->ulevels[t2] - 1 && k < m->nlevels[t2] - 1) {
        const double u = uv[((size_t)t2 * nl + k) * 2], v = uv[((size_t)t2 * nl + k) * 2 + 1];
        vf += (v * cr[2] - u * cr[3]) * helem[(size_t)t2 * nl + k];
    }
    return vf;
}

/* area-weighted mean of the element gradients around `node` on levels [k0, k1) into
 * g[k*4 + cx], g[k*4 + cy] (levels with no wet element keep 0) */
static void node_mean_grad(const loop_mesh* m, const double* grad, int node, int k0, int k1,
                           double* g, int cx, int cy) {
    const int nl = m->nl;
    for (int k = k0; k < k1; ++k) {
        double wsum = 0.0, gx = 0.0, gy = 0.0;
        for (int s = m->nie_off[node]; s < m->nie_off[node + 1]; ++s) {
            const int el = m->nie[s];
            if (k < m->ulevels[el] - 1 || k >= m->nlevels[el] - 1) continue;
            const double w = m->elem_area[el];
            wsum += w;
            gx += grad[((size_t)el * nl + k) * 2] * w;
            gy += grad[((size_t)el * nl + k) * 2 + 1] * w;
        }
        if (wsum > 0.0) {
            g[k * 4 + cx] = gx / wsum;
            g[k * 4 + cy] = gy / wsum;
        }
    }
}

int fct_loops_step(const loop_mesh* m, const double* uv, const double* w, const double* helem,
                   const double* hnode, const double* hnode_new, const double* T,
                   const double* Told, double dt, double* Tnew) {
    const int N = m->nod2D, E = m->elem2D, D = m->edge2D, nl = m->nl;
    const size_t NL = (size_t)N * nl, DL = (size_t)D * nl, EL = (size_t)E * nl;
    double* ab = malloc(NL * sizeof(double));
    double* fh = calloc(DL, sizeof(double));      /* upwind horizontal flux, then adf_h */
    double* fv = calloc(NL, sizeof(double));      /* upwind vertical flux, then adf_v */
    double* lo = calloc(NL, sizeof(double));
    double* grad = calloc(EL * 2, sizeof(double));
    double* eg = malloc(4 * (size_t)nl * sizeof(double));
    double* hi_h = calloc(DL, sizeof(double));
    double* bmax = malloc(NL * sizeof(double));
    double* bmin = malloc(NL * sizeof(double));
    double* emax = malloc(EL * sizeof(double));
    double* emin = malloc(EL * sizeof(double));
    double* cmax = malloc(NL * sizeof(double));
    double* cmin = malloc(NL * sizeof(double));
    double* qmax = calloc(NL, sizeof(double));
    double* qmin = calloc(NL, sizeof(double));
    double* sp = calloc(NL, sizeof(double));
    double* sm = calloc(NL, sizeof(double));
    double* rp = malloc(NL * sizeof(double));
    double* rm = malloc(NL * sizeof(double));
    double* inc = calloc(NL, sizeof(double));
    if (!ab || !fh || !fv || !lo || !grad || !eg || !hi_h || !bmax || !bmin || !emax || !emin ||
        !cmax || !cmin || !qmax || !qmin || !sp || !sm || !rp || !rm || !inc)
        return -1;

    /* node -> (edge, side) lists in increasing 2*edge + side: the edge sums below run per
     * node over this list, which adds each node's contributions in the order of a loop
     * over the edges that adds to the first node and subtracts from the second, and
     * leaves the nodes independent of each other (so they can run in parallel) */
    int* noff = calloc((size_t)N + 1, sizeof(int));
    int* nent = malloc(2 * (size_t)D * sizeof(int));
    if (!noff || !nent) return -1;
    for (int e = 0; e < D; ++e) { noff[m->edges[2 * e] + 1]++; noff[m->edges[2 * e + 1] + 1]++; }
    for (int n = 0; n < N; ++n) noff[n + 1] += noff[n];
    {
        int* pos = malloc(((size_t)N + 1) * sizeof(int));
        if (!pos) return -1;
        memcpy(pos, noff, ((size_t)N + 1) * sizeof(int));
        for (int e = 0; e < D; ++e) {
            nent[pos[m->edges[2 * e]]++] = 2 * e;
            nent[pos[m->edges[2 * e + 1]]++] = 2 * e + 1;
        }
        free(pos);
    }

    /* AB2 field for the high-order fluxes (tracer_adv.py:43-46, 65) */
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < NL; ++i) ab[i] = AB_OLD * Told[i] + AB_NEW * T[i];

    /* first-order upwind fluxes of T: horizontal per edge, vertical per node interface */
#pragma omp parallel for schedule(static)
    for (int e = 0; e < D; ++e) {
        const int a = m->edges[2 * e], b = m->edges[2 * e + 1];
        for (int k = 0; k < nl; ++k) {
            const double vf = edge_vflux(m, uv, helem, e, k);
            if (vf == 0.0) continue;
            const double av = fabs(vf);
            fh[(size_t)e * nl + k] =
                -0.5 * (T[(size_t)a * nl + k] * (vf + av) + T[(size_t)b * nl + k] * (vf - av));
        }
    }
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N; ++n) {
        const int top = m->ulevels_nod2D[n] - 1, bot = m->nlevels_nod2D[n] - 1;
        if (bot <= top) continue;
        const size_t r = (size_t)n * nl;
        fv[r + top] = -w[r + top] * T[r + top] * m->area[r + top];
        for (int k = top + 1; k < bot; ++k) {
            const double wk = w[r + k], aw = fabs(wk);
            fv[r + k] = -0.5 * (T[r + k] * (wk + aw) + T[r + k - 1] * (wk - aw)) * m->area[r + k];
        }
    }

    /* low-order solution: edge divergence (wet levels of either triangle), then the
     * vertical divergence and the thickness update on each wet layer */
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N; ++n) {
        for (int q = noff[n]; q < noff[n + 1]; ++q) {
            const int e = nent[q] >> 1, side = nent[q] & 1;
            const int t1 = m->edge_tri[2 * e], t2 = m->edge_tri[2 * e + 1];
            int kmax = 0;
            if (t1 >= 0 && m->nlevels[t1] - 1 > kmax) kmax = m->nlevels[t1] - 1;
            if (t2 >= 0 && m->nlevels[t2] - 1 > kmax) kmax = m->nlevels[t2] - 1;
            for (int k = 0; k < kmax; ++k) {
                const double f = fh[(size_t)e * nl + k];
                const size_t i = (size_t)n * nl + k;
                lo[i] = side ? lo[i] - f : lo[i] + f;
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N; ++n) {
        const size_t r = (size_t)n * nl;
        for (int k = m->ulevels_nod2D[n] - 1; k < m->nlevels_nod2D[n] - 1; ++k) {
            if (hnode_new[r + k] <= 0.0 || m->areasvol[r + k] <= 0.0) {
                lo[r + k] = 0.0;
                continue;
            }
            lo[r + k] = (T[r + k] * hnode[r + k] +
                         (lo[r + k] + (fv[r + k] - fv[r + k + 1])) * dt / m->areasvol[r + k]) /
                        hnode_new[r + k];
        }
    }

    /* element gradients of T on wet element layers */
#pragma omp parallel for schedule(static)
    for (int el = 0; el < E; ++el) {
        const int* v = m->elem_nodes + 3 * (size_t)el;
        const double* g = m->gradient_sca + 6 * (size_t)el;
        for (int k = m->ulevels[el] - 1; k < m->nlevels[el] - 1; ++k) {
            const double t0 = T[(size_t)v[0] * nl + k], t1 = T[(size_t)v[1] * nl + k],
                         t2 = T[(size_t)v[2] * nl + k];
            grad[((size_t)el * nl + k) * 2] = g[0] * t0 + g[1] * t1 + g[2] * t2;
            grad[((size_t)el * nl + k) * 2 + 1] = g[3] * t0 + g[4] * t1 + g[5] * t2;
        }
    }

    /* MFCT high-order horizontal flux of the AB2 field with the up/down gradients */
#pragma omp parallel
    {
    double* egt = malloc(4 * (size_t)nl * sizeof(double));   /* this thread's slots */
#pragma omp for schedule(static)
    for (int e = 0; e < D; ++e) {
        double* eg = egt;
        const int a = m->edges[2 * e], b = m->edges[2 * e + 1];
        const int up = m->edge_up_dn_tri[2 * e], dn = m->edge_up_dn_tri[2 * e + 1];
        memset(eg, 0, 4 * (size_t)nl * sizeof(double));
        if (up >= 0 && dn >= 0) {
            const int kin0 = (m->ulevels_nod2D_max[a] > m->ulevels_nod2D_max[b]
                                  ? m->ulevels_nod2D_max[a] : m->ulevels_nod2D_max[b]) - 1;
            const int kin1 = (m->nlevels_nod2D_min[a] < m->nlevels_nod2D_min[b]
                                  ? m->nlevels_nod2D_min[a] : m->nlevels_nod2D_min[b]) - 1;
            node_mean_grad(m, grad, a, m->ulevels_nod2D[a] - 1, kin0, eg, 0, 2);
            node_mean_grad(m, grad, b, m->ulevels_nod2D[b] - 1, kin0, eg, 1, 3);
            for (int k = kin0; k < kin1; ++k) {
                eg[k * 4 + 0] = grad[((size_t)up * nl + k) * 2];
                eg[k * 4 + 1] = grad[((size_t)dn * nl + k) * 2];
                eg[k * 4 + 2] = grad[((size_t)up * nl + k) * 2 + 1];
                eg[k * 4 + 3] = grad[((size_t)dn * nl + k) * 2 + 1];
            }
            node_mean_grad(m, grad, a, kin1, m->nlevels_nod2D[a] - 1, eg, 0, 2);
            node_mean_grad(m, grad, b, kin1, m->nlevels_nod2D[b] - 1, eg, 1, 3);
        } else {
            node_mean_grad(m, grad, a, m->ulevels_nod2D[a] - 1, m->nlevels_nod2D[a] - 1, eg, 0, 2);
            node_mean_grad(m, grad, b, m->ulevels_nod2D[b] - 1, m->nlevels_nod2D[b] - 1, eg, 1, 3);
        }
        const int t1 = m->edge_tri[2 * e], t2 = m->edge_tri[2 * e + 1];
        double cosr = t1 >= 0 ? R_EARTH * m->elem_cos[t1] : 0.0;
        if (t2 >= 0) cosr = 0.5 * (cosr + R_EARTH * m->elem_cos[t2]);
        const double ex = m->edge_dxdy[2 * e], ey = m->edge_dxdy[2 * e + 1];
        for (int k = 0; k < nl; ++k) {
            const double vf = edge_vflux(m, uv, helem, e, k);
            if (vf == 0.0) continue;
            const double ta = ab[(size_t)a * nl + k], tb = ab[(size_t)b * nl + k];
            const double d = tb - ta;
            const double* g = eg + 4 * k;
            const double tma = ta + (2 * d + ex * cosr * g[0] + ey * R_EARTH * g[2]) / 6;
            const double tmb = tb - (2 * d + ex * cosr * g[1] + ey * R_EARTH * g[3]) / 6;
            hi_h[(size_t)e * nl + k] = -0.5 * ((vf + fabs(vf)) * tma + (vf - fabs(vf)) * tmb);
        }
    }
    free(egt);
    }
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < DL; ++i) fh[i] = hi_h[i] - fh[i];   /* antidiffusive adf_h */

    /* QR4C high-order vertical flux of the AB2 field minus the upwind flux: adf_v */
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N; ++n) {
        const int top = m->ulevels_nod2D[n] - 1, bot = m->nlevels_nod2D[n] - 1;
        if (bot <= top) continue;
        const size_t r = (size_t)n * nl;
        const double* A = ab + r;
        const double* W = w + r;
        const double* ar = m->area + r;
        fv[r + top] = -A[top] * W[top] * ar[top] - fv[r + top];
        if (bot - top >= 2) {
            fv[r + top + 1] = -0.5 * (A[top] + A[top + 1]) * W[top + 1] * ar[top + 1] - fv[r + top + 1];
            if (bot - 1 != top + 1)
                fv[r + bot - 1] =
                    -0.5 * (A[bot - 2] + A[bot - 1]) * W[bot - 1] * ar[bot - 1] - fv[r + bot - 1];
        }
        fv[r + bot] = 0.0 - fv[r + bot];
        for (int k = top + 2; k < bot - 1; ++k) {
            const double* Z = m->Z;
            const double qc = (A[k - 1] - A[k]) / (Z[k - 1] - Z[k]);
            const double qu = (A[k] - A[k + 1]) / (Z[k] - Z[k + 1]);
            const double qd = (A[k - 2] - A[k - 1]) / (Z[k - 2] - Z[k - 1]);
            const double tm1 = A[k] + (2 * qc + qu) * (m->zbar[k] - Z[k]) / 3;
            const double tm2 = A[k - 1] + (2 * qc + qd) * (m->zbar[k] - Z[k - 1]) / 3;
            fv[r + k] = -0.5 * (tm1 + tm2) * W[k] * ar[k] - fv[r + k];
        }
    }

    /* Zalesak limiter: bounds from LO and T over the element cluster, sums of
     * incoming/outgoing antidiffusive fluxes, ratios, limited fluxes */
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < NL; ++i) {
        bmax[i] = dmax(lo[i], T[i]);
        bmin[i] = dmin(lo[i], T[i]);
    }
    for (size_t i = 0; i < EL; ++i) { emax[i] = -BIG; emin[i] = BIG; }
#pragma omp parallel for schedule(static)
    for (int el = 0; el < E; ++el) {
        const int* v = m->elem_nodes + 3 * (size_t)el;
        for (int k = m->ulevels[el] - 1; k < m->nlevels[el] - 1; ++k) {
            const size_t i0 = (size_t)v[0] * nl + k, i1 = (size_t)v[1] * nl + k,
                         i2 = (size_t)v[2] * nl + k;
            emax[(size_t)el * nl + k] = dmax(dmax(bmax[i0], bmax[i1]), bmax[i2]);
            emin[(size_t)el * nl + k] = dmin(dmin(bmin[i0], bmin[i1]), bmin[i2]);
        }
    }
    for (size_t i = 0; i < NL; ++i) { cmax[i] = -BIG; cmin[i] = BIG; }
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N; ++n) {
        const size_t r = (size_t)n * nl;
        for (int k = m->ulevels_nod2D[n] - 1; k < m->nlevels_nod2D[n] - 1; ++k)
            for (int s = m->nie_off[n]; s < m->nie_off[n + 1]; ++s) {
                const size_t ie = (size_t)m->nie[s] * nl + k;
                cmax[r + k] = dmax(cmax[r + k], emax[ie]);
                cmin[r + k] = dmin(cmin[r + k], emin[ie]);
            }
    }
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N; ++n) {
        const int top = m->ulevels_nod2D[n] - 1, bot = m->nlevels_nod2D[n] - 1;
        const size_t r = (size_t)n * nl;
        for (int k = top; k < bot; ++k) {
            if (k == top || k == bot - 1) {
                qmax[r + k] = cmax[r + k] - lo[r + k];
                qmin[r + k] = cmin[r + k] - lo[r + k];
            } else {
                qmax[r + k] = dmax(dmax(cmax[r + k - 1], cmax[r + k]), cmax[r + k + 1]) - lo[r + k];
                qmin[r + k] = dmin(dmin(cmin[r + k - 1], cmin[r + k]), cmin[r + k + 1]) - lo[r + k];
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N; ++n) {
        const size_t r = (size_t)n * nl;
        for (int k = m->ulevels_nod2D[n] - 1; k < m->nlevels_nod2D[n] - 1; ++k) {
            const double ft = fv[r + k], fb = fv[r + k + 1];
            sp[r + k] += dmax(ft, 0.0) + dmax(-fb, 0.0);
            sm[r + k] += dmin(ft, 0.0) + dmin(-fb, 0.0);
        }
    }
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N; ++n) {
        const size_t r = (size_t)n * nl;
        for (int q = noff[n]; q < noff[n + 1]; ++q) {
            const int e = nent[q] >> 1, side = nent[q] & 1;
            for (int k = 0; k < nl; ++k) {
                const double f = fh[(size_t)e * nl + k];
                const double g = side ? -f : f;
                sp[r + k] += dmax(g, 0.0);
                sm[r + k] += dmin(g, 0.0);
            }
        }
    }
    for (size_t i = 0; i < NL; ++i) { rp[i] = 1.0; rm[i] = 1.0; }
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N; ++n) {
        const size_t r = (size_t)n * nl;
        for (int k = m->ulevels_nod2D[n] - 1; k < m->nlevels_nod2D[n] - 1; ++k) {
            const double av = m->areasvol[r + k], hn = hnode_new[r + k];
            if (av <= 0.0 || hn <= 0.0) continue;
            rp[r + k] = dmin(1.0, qmax[r + k] / (sp[r + k] * dt / av / hn + FLUX_EPS));
            rm[r + k] = dmin(1.0, qmin[r + k] / (sm[r + k] * dt / av / hn - FLUX_EPS));
        }
    }
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N; ++n) {
        const int top = m->ulevels_nod2D[n] - 1, bot = m->nlevels_nod2D[n] - 1;
        if (bot <= top) continue;
        const size_t r = (size_t)n * nl;
        double f = fv[r + top];
        fv[r + top] = (f >= 0.0 ? rp[r + top] : rm[r + top]) * f;
        for (int k = top + 1; k < bot; ++k) {
            f = fv[r + k];
            const double lim = f >= 0.0 ? dmin(rm[r + k - 1], rp[r + k]) : dmin(rp[r + k - 1], rm[r + k]);
            fv[r + k] = lim * f;
        }
    }
#pragma omp parallel for schedule(static)
    for (int e = 0; e < D; ++e) {
        const size_t ra = (size_t)m->edges[2 * e] * nl, rb = (size_t)m->edges[2 * e + 1] * nl;
        for (int k = 0; k < nl; ++k) {
            const double f = fh[(size_t)e * nl + k];
            const double lim = f >= 0.0 ? dmin(rp[ra + k], rm[rb + k]) : dmin(rm[ra + k], rp[rb + k]);
            fh[(size_t)e * nl + k] = lim * f;
        }
    }

    /* LO plus the limited antidiffusive divergence, then the ALE reconstruction */
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N; ++n) {
        const size_t r = (size_t)n * nl;
        for (int k = m->ulevels_nod2D[n] - 1; k < m->nlevels_nod2D[n] - 1; ++k) {
            inc[r + k] += -T[r + k] * hnode[r + k] + lo[r + k] * hnode_new[r + k];
            if (m->areasvol[r + k] > 0.0)
                inc[r + k] += (fv[r + k] - fv[r + k + 1]) * dt / m->areasvol[r + k];
        }
    }
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N; ++n) {
        const size_t r = (size_t)n * nl;
        for (int q = noff[n]; q < noff[n + 1]; ++q) {
            const int e = nent[q] >> 1, side = nent[q] & 1;
            for (int k = 0; k < nl; ++k) {
                if (m->areasvol[r + k] <= 0.0) continue;
                const double d = fh[(size_t)e * nl + k] * dt / m->areasvol[r + k];
                inc[r + k] = side ? inc[r + k] - d : inc[r + k] + d;
            }
        }
    }
    memcpy(Tnew, T, NL * sizeof(double));
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N; ++n) {
        const size_t r = (size_t)n * nl;
        for (int k = m->ulevels_nod2D[n] - 1; k < m->nlevels_nod2D[n] - 1; ++k) {
            inc[r + k] += T[r + k] * (hnode[r + k] - hnode_new[r + k]);
            if (hnode_new[r + k] > 0.0) Tnew[r + k] += inc[r + k] / hnode_new[r + k];
        }
    }

    free(ab); free(fh); free(fv); free(lo); free(grad); free(eg); free(hi_h);
    free(bmax); free(bmin); free(emax); free(emin); free(cmax); free(cmin);
    free(qmax); free(qmin); free(sp); free(sm); free(rp); free(rm); free(inc);
    free(noff); free(nent);
    return 0;
}

/* Vertical velocity by continuity with the original C kernel's loops
 * (fesom_ale_vert_vel_linfs as described at fesom_jax/ale.py:309-336): per edge and
 * wet triangle level, c = (v dx - u dy) h of the left triangle minus that of the right
 * one, added to the first node and subtracted from the second; then, per node from the
 * level above the bottom interface up to the surface, w[k] += w[k+1]; then / area where
 * area > 0. Interfaces outside [nzmin, nzmax] are zero. */
int loops_compute_w(const loop_mesh* m, const double* uv, const double* helem, double* w) {
    const int N = m->nod2D, D = m->edge2D, nl = m->nl;
    memset(w, 0, (size_t)N * nl * sizeof(double));
    for (int e = 0; e < D; ++e) {
        const int a = m->edges[2 * e], b = m->edges[2 * e + 1];
        const int t1 = m->edge_tri[2 * e], t2 = m->edge_tri[2 * e + 1];
        const double* cr = m->edge_cross_dxdy + 4 * (size_t)e;
        for (int k = 0; k < nl; ++k) {
            double c1 = 0.0, c2 = 0.0;
            if (t1 >= 0 && k >= m->ulevels[t1] - 1 && k < m->nlevels[t1] - 1) {
                const size_t i = (size_t)t1 * nl + k;
                c1 = (uv[2 * i + 1] * cr[0] - uv[2 * i] * cr[1]) * helem[i];
            }
            if (t2 >= 0 && k >= m->ulevels[t2] - 1 && k < m->nlevels[t2] - 1) {
                const size_t i = (size_t)t2 * nl + k;
                c2 = (uv[2 * i + 1] * cr[2] - uv[2 * i] * cr[3]) * helem[i];
            }
            const double c = c1 - c2;
            w[(size_t)a * nl + k] += c;
            w[(size_t)b * nl + k] -= c;
        }
    }
    for (int n = 0; n < N; ++n) {
        const int top = m->ulevels_nod2D[n] - 1, bot = m->nlevels_nod2D[n] - 1;
        double* col = w + (size_t)n * nl;
        for (int k = bot - 1; k >= top; --k) col[k] += col[k + 1];
        for (int k = 0; k < nl; ++k) {
            if (k < top || k > bot) { col[k] = 0.0; continue; }
            const double ar = m->area[(size_t)n * nl + k];
            if (ar > 0.0) col[k] = col[k] / ar;
        }
    }
    return 0;
}

/* Sea-ice FCT of one tracer set with element loops and scatter-adds, the structure of
 * the original fesom_ice_fct.c as fesom_jax/ice_adv.py:1-32, 111-194 describes it:
 * Taylor-Galerkin element matrices, consistent-mass low/high-order solves, element
 * antidiffusive fluxes, node bounds over the surrounding elements, node ratios, element
 * limiting by the minimum over its vertices. Elements or nodes under an ice shelf
 * (ulevels > 1) take no part. */
typedef struct {
    double dt, diff, gamma;
} loop_ice_cfg;

static void ice_mm(const loop_mesh* m, const double* X, double* out) {
    memset(out, 0, (size_t)m->nod2D * sizeof(double));
    for (int e = 0; e < m->elem2D; ++e) {
        if (m->ulevels[e] > 1) continue;
        const int* v = m->elem_nodes + 3 * (size_t)e;
        const double s = (X[v[0]] + X[v[1]]) + X[v[2]];
        const double w = m->elem_area[e] / 12.0;
        for (int n = 0; n < 3; ++n) out[v[n]] += w * (X[v[n]] + s);
    }
}

static void ice_one(const loop_mesh* m, const loop_ice_cfg* c, const double* ent,
                    const double* X, double* out) {
    const int N = m->nod2D, E = m->elem2D, nl = m->nl;
    double* rhs = calloc(N, sizeof(double));
    double* mm = malloc(N * sizeof(double));
    double* xl = malloc(N * sizeof(double));
    double* d = malloc(N * sizeof(double));
    double* w = malloc(N * sizeof(double));
    double* hi = malloc(N * sizeof(double));
    double* lo = malloc(N * sizeof(double));
    double* fp = calloc(N, sizeof(double));
    double* fm = calloc(N, sizeof(double));
    double* pp = malloc(N * sizeof(double));
    double* pm = malloc(N * sizeof(double));
    double* acc = calloc(N, sizeof(double));
    double* flux = malloc(3 * (size_t)E * sizeof(double));
    /* right-hand side */
    for (int e = 0; e < E; ++e) {
        if (m->ulevels[e] > 1) continue;
        const int* v = m->elem_nodes + 3 * (size_t)e;
        const double* A = ent + 9 * (size_t)e;
        for (int n = 0; n < 3; ++n)
            rhs[v[n]] += (A[3 * n] * X[v[0]] + A[3 * n + 1] * X[v[1]]) + A[3 * n + 2] * X[v[2]];
    }
    /* low order: (rhs + gamma mm X) / area + (1 - gamma) X */
    ice_mm(m, X, mm);
    for (int n = 0; n < N; ++n) {
        const double ar = m->area[(size_t)n * nl];
        const double inv = 1.0 / (ar > 0 ? ar : 1.0);
        xl[n] = m->ulevels_nod2D[n] <= 1 ? (rhs[n] + c->gamma * mm[n]) * inv + (1.0 - c->gamma) * X[n]
                                         : X[n];
        d[n] = m->ulevels_nod2D[n] <= 1 ? rhs[n] * inv : 0.0;
    }
    /* high order: two residual corrections with the consistent mass */
    for (int it = 0; it < 2; ++it) {
        ice_mm(m, d, mm);
        for (int n = 0; n < N; ++n) {
            if (m->ulevels_nod2D[n] > 1) continue;
            const double ar = m->area[(size_t)n * nl];
            const double inv = 1.0 / (ar > 0 ? ar : 1.0);
            d[n] = d[n] + (rhs[n] - mm[n]) * inv;
        }
    }
    /* antidiffusive element fluxes and node bounds */
    for (int n = 0; n < N; ++n) {
        w[n] = c->gamma * X[n] + d[n];
        hi[n] = -INFINITY;
        lo[n] = INFINITY;
    }
    for (int e = 0; e < E; ++e) {
        const int* v = m->elem_nodes + 3 * (size_t)e;
        for (int q = 0; q < 3; ++q) flux[3 * (size_t)e + q] = 0.0;
        if (m->ulevels[e] > 1) continue;
        const double s = (w[v[0]] + w[v[1]]) + w[v[2]];
        double eh = -INFINITY, el = INFINITY;
        for (int q = 0; q < 3; ++q) {
            const double ar = m->area[(size_t)v[q] * nl];
            const double inv = 1.0 / (ar > 0 ? ar : 1.0);
            const double f = -(s - 3.0 * w[v[q]]) * (m->elem_area[e] * inv) / 12.0;
            flux[3 * (size_t)e + q] = f;
            fp[v[q]] += f > 0.0 ? f : 0.0;
            fm[v[q]] += f < 0.0 ? f : 0.0;
            eh = dmax(eh, dmax(xl[v[q]], X[v[q]]));
            el = dmin(el, dmin(xl[v[q]], X[v[q]]));
        }
        for (int q = 0; q < 3; ++q) {
            hi[v[q]] = dmax(hi[v[q]], eh);
            lo[v[q]] = dmin(lo[v[q]], el);
        }
    }
    for (int n = 0; n < N; ++n) {
        pp[n] = 0.0;
        pm[n] = 0.0;
        if (m->ulevels_nod2D[n] > 1) continue;
        const double h = isfinite(hi[n]) ? hi[n] : 0.0, l = isfinite(lo[n]) ? lo[n] : 0.0;
        if (fp[n] > 0.0) pp[n] = dmin((h - xl[n]) / (fp[n] > 1e-12 ? fp[n] : 1e-12), 1.0);
        if (fm[n] < 0.0) pm[n] = dmin((l - xl[n]) / (fm[n] < -1e-12 ? fm[n] : -1e-12), 1.0);
    }
    /* limit each element's fluxes by the smallest vertex ratio and scatter them */
    for (int e = 0; e < E; ++e) {
        if (m->ulevels[e] > 1) continue;
        const int* v = m->elem_nodes + 3 * (size_t)e;
        const double* f = flux + 3 * (size_t)e;
        double ae = INFINITY;
        for (int q = 0; q < 3; ++q) ae = dmin(ae, f[q] >= 0.0 ? pp[v[q]] : pm[v[q]]);
        for (int q = 0; q < 3; ++q) acc[v[q]] += f[q] * ae;
    }
    for (int n = 0; n < N; ++n) out[n] = xl[n] + acc[n];
    free(rhs); free(mm); free(xl); free(d); free(w); free(hi); free(lo); free(fp); free(fm);
    free(pp); free(pm); free(acc); free(flux);
}

int loops_ice_fct(const loop_mesh* m, const loop_ice_cfg* c, const double* a_ice,
                  const double* m_ice, const double* m_snow, const double* u_ice,
                  const double* v_ice, double* a_out, double* m_out, double* ms_out) {
    const int E = m->elem2D;
    double* ent = malloc(9 * (size_t)E * sizeof(double));
    if (!ent) return -1;
    /* Taylor-Galerkin element matrices vol dt (a - b - c), rows n, columns q */
    for (int e = 0; e < E; ++e) {
        const int* v = m->elem_nodes + 3 * (size_t)e;
        const double* g = m->gradient_sca + 6 * (size_t)e;
        const double vol = m->elem_area[e];
        const double U[3] = {u_ice[v[0]], u_ice[v[1]], u_ice[v[2]]};
        const double V[3] = {v_ice[v[0]], v_ice[v[1]], v_ice[v[2]]};
        const double um = (U[0] + U[1]) + U[2], vm = (V[0] + V[1]) + V[2];
        const double diff = c->diff * sqrt(vol / 2.0e8);
        double udn[3];
        for (int n = 0; n < 3; ++n) udn[n] = um * g[n] + vm * g[3 + n];
        for (int n = 0; n < 3; ++n)
            for (int q = 0; q < 3; ++q) {
                const double a = (g[n] * (um + U[q]) + g[3 + n] * (vm + V[q])) / 12.0;
                const double b = diff * (g[n] * g[q] + g[3 + n] * g[3 + q]);
                const double cc = 0.5 * c->dt * (udn[n] * udn[q]) / 9.0;
                ent[9 * (size_t)e + 3 * n + q] = vol * c->dt * (a - b - cc);
            }
    }
    ice_one(m, c, ent, a_ice, a_out);
    ice_one(m, c, ent, m_ice, m_out);
    ice_one(m, c, ent, m_snow, ms_out);
    free(ent);
    return 0;
}
