// Sea-ice FCT advection of the three ice tracers (fct_solve, fesom_jax/ice_adv.py:37-194):
// Taylor-Galerkin right-hand side, consistent-mass low- and high-order solves, element
// antidiffusive fluxes and the Zalesak limiter, on the surface nodes.
//
// Mapping: one thread per node and all three tracers per thread (the problem is 2-D;
// the tracers share every geometry load). Each node gathers over its elements in
// increasing element id (the node -> element CSR), which is the order of the
// reference's scatter-add over (element, vertex); sums over an element's three
// vertices run ((x0 + x1) + x2). Six launches: the inputs packed per node, then five
// gather passes over the elements around every node: right-hand side + low-order
// solve, two high-order corrections (the second also stores each node's gamma * X + d,
// the vertex value of the antidiffusive fluxes), fluxes + bounds + ratios, limited
// fluxes + update.
//
// Layout: every gather is per-lane scattered, so the passes are bound by the number
// of gather instructions. The three tracers of a node sit together (dv4, the fourth
// slot carrying the node's inverse area where a pass needs it), (u, v) as a pair, and
// an element's vertices, open flag, gradient coefficients, volume and diffusion factor
// in one iv4 and four dv2: a vertex costs two loads for all three tracers instead
// of three, an element five instead of twelve.

#include <hip/hip_runtime.h>

#include "fct_common.h"

namespace fct {
namespace {

template <class T>
using gp = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gp<T> glb(T* p) { return (gp<T>)p; }

// packed per-node and per-element values as clang vectors (HIP's vector types are
// classes, which do not load through address-space pointers)
typedef double dv2 __attribute__((ext_vector_type(2)));
typedef double dv4 __attribute__((ext_vector_type(4)));
typedef int iv4 __attribute__((ext_vector_type(4)));

constexpr double ICE_FLUX_EPS = 1.0e-12;   // ice_adv.py:32
constexpr int ICE_K = 3;                   // a_ice, m_ice, m_snow

struct IceArgs {
    int N;
    double dt, diff, gamma;
    const int* ncsr_off;
    const int* ncsr;
    const iv4* el;          // [E] (v0, v1, v2, 1 if the element is open: ulevels <= 1)
    const dv2* eg;          // [E][4] (dx0, dx1), (dx2, dy0), (dy1, dy2), (vol, sqrt(vol/2e8))
    const int* nlev;        // [N*2] (ulevels_nod2D-1, nlevels_nod2D-1)
    const dv4* X;           // [N] (a_ice, m_ice, m_snow, 0), packed from the inputs
    const dv2* uv;          // [N] (u, v)
    double* out[ICE_K];     // tracers out
    dv4* rhs;               // [N] right-hand side
    dv4* xl;                // [N] low-order solution; .w = 1 / surface area
    dv4* d0;                // [N] high-order increment (ping)
    dv4* d1;                // [N] (pong)
    dv4* w;                 // [N] gamma * X + d with the final d; .w = 1 / surface area
    double* pp;             // [3][N] ratios
    double* pm;
};

__device__ __forceinline__ double comp(const dv4& v, int t) {
    return t == 0 ? v.x : (t == 1 ? v.y : v.z);
}
__device__ __forceinline__ bool node_open(const IceArgs& a, int n) {
    return glb(a.nlev)[2 * n] == 0;            // ulevels_nod2D <= 1
}
__device__ __forceinline__ int slot_of(int n, int v0, int v1) { return v0 == n ? 0 : (v1 == n ? 1 : 2); }
__device__ __forceinline__ double sum3(double x0, double x1, double x2) { return (x0 + x1) + x2; }
__device__ __forceinline__ double min1(double x) { return x > 1.0 ? 1.0 : x; }   // jnp.minimum(x, 1)

// The elements of node n in CSR order (increasing id). Each element's table entry is
// fetched one element ahead and its CSR entry two ahead, so that an element's vertex
// gathers need not wait behind the chain CSR entry -> element entry.
struct ElemWalk {
    int s, s1, e_next, e_after;
    iv4 ev_next;
    __device__ __forceinline__ ElemWalk(const IceArgs& a, int n) {
        s = glb(a.ncsr_off)[n];
        s1 = glb(a.ncsr_off)[n + 1];
        e_next = s < s1 ? glb(a.ncsr)[s] : 0;
        e_after = s + 1 < s1 ? glb(a.ncsr)[s + 1] : 0;
        ev_next = s < s1 ? glb(a.el)[e_next] : iv4{0, 0, 0, 0};
    }
    __device__ __forceinline__ bool next(const IceArgs& a, int& e, iv4& ev) {
        if (s >= s1) return false;
        e = e_next;
        ev = ev_next;
        e_next = e_after;
        if (s + 1 < s1) ev_next = glb(a.el)[e_next];
        if (s + 2 < s1) e_after = glb(a.ncsr)[s + 2];
        ++s;
        return true;
    }
};

// (mm X)[n]'s term of element e: vol/12 (X[n] + sum of the vertices)
__device__ __forceinline__ double mm_term(double vol, double xq, double x0, double x1, double x2) {
    return vol / 12.0 * (xq + sum3(x0, x1, x2));
}

// element tables, once per plan: vertices and open flag; gradient coefficients, volume
// and the diffusion factor sq (computed on the host)
__global__ void __launch_bounds__(256) k_ice_tables(int E, const int* __restrict__ en,
                                                    const int2* __restrict__ elev,
                                                    const double* __restrict__ gs,
                                                    const double* __restrict__ vol,
                                                    const double* __restrict__ sq,
                                                    iv4* __restrict__ el, dv2* __restrict__ eg) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    el[e] = iv4{en[3 * e], en[3 * e + 1], en[3 * e + 2], elev[e].x == 0 ? 1 : 0};
    eg[4 * e] = dv2{gs[6 * e], gs[6 * e + 1]};
    eg[4 * e + 1] = dv2{gs[6 * e + 2], gs[6 * e + 3]};
    eg[4 * e + 2] = dv2{gs[6 * e + 4], gs[6 * e + 5]};
    eg[4 * e + 3] = dv2{vol[e], sq[e]};
}

// the inputs packed per node
__global__ void __launch_bounds__(256) k_ice_pack(int N, const double* __restrict__ a0,
                                                  const double* __restrict__ a1,
                                                  const double* __restrict__ a2,
                                                  const double* __restrict__ u,
                                                  const double* __restrict__ v,
                                                  dv4* __restrict__ X, dv2* __restrict__ uv) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    X[n] = dv4{a0[n], a1[n], a2[n], 0.0};
    uv[n] = dv2{u[n], v[n]};
}

// pass 1: right-hand side (ice_adv.py:47-73), low-order solution (:76-87) and the first
// high-order approximation rhs / area (:101)
__global__ void __launch_bounds__(256) k_ice_rhs(IceArgs a, const double* __restrict__ area,
                                                 int L) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.N) return;
    const auto UV = glb(a.uv);
    const auto X = glb(a.X);
    double rhs[ICE_K] = {0.0, 0.0, 0.0}, mm[ICE_K] = {0.0, 0.0, 0.0};
    const dv4 xn = X[n];
    ElemWalk wk(a, n);
    int e;
    iv4 ev;
    while (wk.next(a, e, ev)) {
        const int vv[3] = {ev.x, ev.y, ev.z};
        const int q = slot_of(n, vv[0], vv[1]);
        const bool on = ev.w != 0;
        const dv2 g0 = glb(a.eg)[4 * e], g1 = glb(a.eg)[4 * e + 1];
        const dv2 g2 = glb(a.eg)[4 * e + 2], g3 = glb(a.eg)[4 * e + 3];
        const double dx[3] = {g0.x, g0.y, g1.x}, dy[3] = {g1.y, g2.x, g2.y};
        const double vol = g3.x;
        double Uv[3], Vv[3];
        dv4 xv[3];
        for (int i = 0; i < 3; ++i) {
            const dv2 w = UV[vv[i]];
            Uv[i] = w.x;
            Vv[i] = w.y;
            xv[i] = X[vv[i]];
        }
        const double um = sum3(Uv[0], Uv[1], Uv[2]), vm = sum3(Vv[0], Vv[1], Vv[2]);
        const double diff = a.diff * g3.y;
        const double udn_q = um * dx[q] + vm * dy[q];
        double ent[3];
        for (int j = 0; j < 3; ++j) {
            const double ta = (dx[q] * (um + Uv[j]) + dy[q] * (vm + Vv[j])) / 12.0;
            const double tb = diff * (dx[q] * dx[j] + dy[q] * dy[j]);
            const double udn_j = um * dx[j] + vm * dy[j];
            const double tc = 0.5 * a.dt * (udn_q * udn_j) / 9.0;
            ent[j] = vol * a.dt * (ta - tb - tc);
        }
        for (int t = 0; t < ICE_K; ++t) {
            const double x0 = comp(xv[0], t), x1 = comp(xv[1], t), x2 = comp(xv[2], t);
            const double c = (ent[0] * x0 + ent[1] * x1) + ent[2] * x2;
            const double xq = q == 0 ? x0 : (q == 1 ? x1 : x2);
            rhs[t] = rhs[t] + (on ? c : 0.0);
            mm[t] = mm[t] + (on ? mm_term(vol, xq, x0, x1, x2) : 0.0);
        }
    }
    const double ar = glb(area)[(size_t)n * L];
    const double inv = 1.0 / (ar > 0 ? ar : 1.0);
    const bool nc = node_open(a, n);
    double xl[ICE_K], d[ICE_K];
    for (int t = 0; t < ICE_K; ++t) {
        const double x = comp(xn, t);
        xl[t] = nc ? (rhs[t] + a.gamma * mm[t]) * inv + (1.0 - a.gamma) * x : x;
        d[t] = nc ? rhs[t] * inv : 0.0;
    }
    glb(a.rhs)[n] = dv4{rhs[0], rhs[1], rhs[2], 0.0};
    glb(a.xl)[n] = dv4{xl[0], xl[1], xl[2], inv};
    glb(a.d0)[n] = dv4{d[0], d[1], d[2], 0.0};
}

// pass 2 (twice): d <- d + (rhs - mm d) / area (ice_adv.py:102-104); the second pass also
// stores gamma * X + d (wout), the vertex value of the antidiffusive fluxes
__global__ void __launch_bounds__(256) k_ice_hi(IceArgs a, const dv4* din, dv4* dout,
                                                dv4* wout) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.N) return;
    const auto D = glb(din);
    double mm[ICE_K] = {0.0, 0.0, 0.0};
    // one element ahead: its area and vertex values load while this one is summed
    ElemWalk wk(a, n);
    int e;
    iv4 ev;
    bool more = wk.next(a, e, ev);
    double vol_n = 0.0;
    dv4 dn0 = {}, dn1 = {}, dn2 = {};
    if (more) {
        vol_n = glb(a.eg)[4 * e + 3].x;
        dn0 = D[ev.x]; dn1 = D[ev.y]; dn2 = D[ev.z];
    }
    while (more) {
        const int q = slot_of(n, ev.x, ev.y);
        const bool on = ev.w != 0;
        const double vol = vol_n;
        const dv4 d0 = dn0, d1 = dn1, d2 = dn2;
        more = wk.next(a, e, ev);
        if (more) {
            vol_n = glb(a.eg)[4 * e + 3].x;
            dn0 = D[ev.x]; dn1 = D[ev.y]; dn2 = D[ev.z];
        }
        for (int t = 0; t < ICE_K; ++t) {
            const double x0 = comp(d0, t), x1 = comp(d1, t), x2 = comp(d2, t);
            const double xq = q == 0 ? x0 : (q == 1 ? x1 : x2);
            mm[t] = mm[t] + (on ? mm_term(vol, xq, x0, x1, x2) : 0.0);
        }
    }
    const dv4 xl = glb(a.xl)[n];
    const double inv = xl.w;
    const bool nc = node_open(a, n);
    const dv4 dv = D[n], r = glb(a.rhs)[n];
    double dn[ICE_K];
    for (int t = 0; t < ICE_K; ++t) {
        const double d = comp(dv, t);
        dn[t] = nc ? d + (comp(r, t) - mm[t]) * inv : d;
    }
    glb(dout)[n] = dv4{dn[0], dn[1], dn[2], 0.0};
    if (wout) {
        const dv4 x = glb(a.X)[n];
        glb(wout)[n] = dv4{a.gamma * x.x + dn[0], a.gamma * x.y + dn[1],
                           a.gamma * x.z + dn[2], inv};
    }
}

// antidiffusive flux of element e at its vertex q for one tracer (ice_adv.py:125-130):
// ice_flux = ice_flux_num / 12
__device__ __forceinline__ double ice_flux_num(const double (&w)[3], int q, double vol,
                                               double inv_q) {
    const double s = sum3(w[0], w[1], w[2]) - 3.0 * w[q];
    return -s * (vol * inv_q);
}
__device__ __forceinline__ double ice_flux(double g, const double (&w)[3], int q, double vol,
                                           double inv_q) {
    return ice_flux_num(w, q, vol, inv_q) / 12.0;
}
// ice_flux(...) >= 0.0 from its numerator g, without the division: g / 12 rounds to a
// value >= 0 (or to -0) exactly when g >= -6 * 2^-1074 (|g / 12| <= 2^-1075 rounds to
// zero, ties to even); NaN compares false either way
__device__ __forceinline__ bool ice_flux_nonneg(double g) { return g >= -0x6p-1074; }

// pass 3: flux sums, cluster bounds and clipping ratios of every node (ice_adv.py:132-160)
__global__ void __launch_bounds__(256) k_ice_ratio(IceArgs a) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.N) return;
    const auto W = glb(a.w);
    const auto XL = glb(a.xl);
    const auto X = glb(a.X);
    const dv4 xln = XL[n];
    const double inv = xln.w;
    double fp[ICE_K] = {0.0, 0.0, 0.0}, fm[ICE_K] = {0.0, 0.0, 0.0};
    double hi[ICE_K], lo[ICE_K];
    for (int t = 0; t < ICE_K; ++t) { hi[t] = -INFINITY; lo[t] = INFINITY; }
    ElemWalk wk(a, n);
    int e;
    iv4 ev;
    while (wk.next(a, e, ev)) {
        const int vv[3] = {ev.x, ev.y, ev.z};
        const int q = slot_of(n, vv[0], vv[1]);
        const bool on = ev.w != 0;
        const double vol = glb(a.eg)[4 * e + 3].x;
        dv4 xv[3], lv[3], wv[3];
        for (int j = 0; j < 3; ++j) {
            xv[j] = X[vv[j]];
            lv[j] = XL[vv[j]];
            wv[j] = W[vv[j]];
        }
        for (int t = 0; t < ICE_K; ++t) {
            double w[3], eh = -INFINITY, el = INFINITY;
            for (int j = 0; j < 3; ++j) {
                const double x = comp(xv[j], t), l = comp(lv[j], t);
                w[j] = comp(wv[j], t);
                eh = fmax(eh, fmax(l, x));
                el = fmin(el, fmin(l, x));
            }
            const double f = on ? ice_flux(a.gamma, w, q, vol, inv) : 0.0;
            fp[t] = fp[t] + fmax(f, 0.0);
            fm[t] = fm[t] + fmin(f, 0.0);
            if (on) {
                hi[t] = fmax(hi[t], eh);
                lo[t] = fmin(lo[t], el);
            }
        }
    }
    const bool nc = node_open(a, n);
    for (int t = 0; t < ICE_K; ++t) {
        const size_t i = (size_t)t * a.N + n;
        const double h = isfinite(hi[t]) ? hi[t] : 0.0, l = isfinite(lo[t]) ? lo[t] : 0.0;
        const double xl = comp(xln, t);
        const double tmax = h - xl, tmin = l - xl;
        double p = fp[t] > 0.0 ? min1(tmax / (fp[t] > ICE_FLUX_EPS ? fp[t] : ICE_FLUX_EPS)) : 0.0;
        double m = fm[t] < 0.0 ? min1(tmin / (fm[t] < -ICE_FLUX_EPS ? fm[t] : -ICE_FLUX_EPS)) : 0.0;
        glb(a.pp)[i] = nc ? p : 0.0;
        glb(a.pm)[i] = nc ? m : 0.0;
    }
}

// pass 4: each element's fluxes limited by the smallest ratio of its vertices, summed
// into the node, added to the low-order solution (ice_adv.py:162-169)
__global__ void __launch_bounds__(256) k_ice_apply(IceArgs a) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.N) return;
    const auto W = glb(a.w);
    const auto PP = glb(a.pp), PM = glb(a.pm);
    double acc[ICE_K] = {0.0, 0.0, 0.0};
    // one element ahead: its area and vertex values load while this one is summed
    ElemWalk wk(a, n);
    int e;
    iv4 ev;
    bool more = wk.next(a, e, ev);
    double vol_n = 0.0;
    dv4 wn[3] = {};
    if (more) {
        vol_n = glb(a.eg)[4 * e + 3].x;
        wn[0] = W[ev.x]; wn[1] = W[ev.y]; wn[2] = W[ev.z];
    }
    while (more) {
        const int vv[3] = {ev.x, ev.y, ev.z};
        const int q = slot_of(n, vv[0], vv[1]);
        const bool on = ev.w != 0;
        const double vol = vol_n;
        // the vertices' gamma * X + d, with their inverse areas in the fourth slot
        const dv4 wv[3] = {wn[0], wn[1], wn[2]};
        more = wk.next(a, e, ev);
        if (more) {
            vol_n = glb(a.eg)[4 * e + 3].x;
            wn[0] = W[ev.x]; wn[1] = W[ev.y]; wn[2] = W[ev.z];
        }
        for (int t = 0; t < ICE_K; ++t) {
            const size_t o = (size_t)t * a.N;
            double w[3];
            for (int j = 0; j < 3; ++j) w[j] = comp(wv[j], t);
            // the element's factor takes each vertex's ratio by the sign of that vertex's
            // flux (0 when the element is closed); only the node's own flux is divided
            double ae = INFINITY, gq = 0.0;
            for (int j = 0; j < 3; ++j) {
                const double g = ice_flux_num(w, j, vol, wv[j].w);
                gq = j == q ? g : gq;
                ae = fmin(ae, !on || ice_flux_nonneg(g) ? PP[o + vv[j]] : PM[o + vv[j]]);
            }
            acc[t] = acc[t] + (on ? (gq / 12.0) * ae : 0.0);
        }
    }
    const dv4 xl = glb(a.xl)[n];
    for (int t = 0; t < ICE_K; ++t) glb(a.out[t])[n] = comp(xl, t) + acc[t];
}

}  // namespace

// rhs, xl, d0, d1, w, X as dv4, (u, v) as dv2, pp and pm: 32 doubles per node
size_t ice_workspace_doubles(int N) { return (size_t)32 * N; }
// el (iv4) and eg (four dv2): 10 doubles per element
size_t ice_table_doubles(int E) { return (size_t)10 * E; }

int launch_ice_tables(const DevMesh& m, const double* sq, double* tab, hipStream_t s) {
    if (m.E == 0) return FCT_OK;
    iv4* el = reinterpret_cast<iv4*>(tab);
    dv2* eg = reinterpret_cast<dv2*>(tab + 2 * (size_t)m.E);
    k_ice_tables<<<(m.E + 255) / 256, 256, 0, s>>>(m.E, m.en, m.elev, m.gs, m.earea, sq, el, eg);
    return check_launch("k_ice_tables");
}

int launch_ice_fct(const DevMesh& m, const IceCall& c, double* work, const double* tab,
                   hipStream_t s) {
    if (m.N == 0) return FCT_OK;
    IceArgs a{};
    a.N = m.N;
    a.dt = c.dt;
    a.diff = c.diff;
    a.gamma = c.gamma;
    a.ncsr_off = m.ncsr_off;
    a.ncsr = m.ncsr;
    a.el = reinterpret_cast<const iv4*>(tab);
    a.eg = reinterpret_cast<const dv2*>(tab + 2 * (size_t)m.E);
    a.nlev = reinterpret_cast<const int*>(m.nlev);
    for (int t = 0; t < ICE_K; ++t) a.out[t] = c.out[t];
    const size_t N = (size_t)m.N;
    dv4* w4 = reinterpret_cast<dv4*>(work);
    a.rhs = w4;
    a.xl = w4 + N;
    a.d0 = w4 + 2 * N;
    a.d1 = w4 + 3 * N;
    a.w = w4 + 4 * N;
    dv4* X = w4 + 5 * N;
    dv2* uv = reinterpret_cast<dv2*>(work + 24 * N);
    a.X = X;
    a.uv = uv;
    a.pp = work + 26 * N;
    a.pm = work + 29 * N;
    const dim3 grid((m.N + 255) / 256), block(256);
    k_ice_pack<<<grid, block, 0, s>>>(m.N, c.X[0], c.X[1], c.X[2], c.u, c.v, X, uv);
    k_ice_rhs<<<grid, block, 0, s>>>(a, m.area, m.L);
    k_ice_hi<<<grid, block, 0, s>>>(a, a.d0, a.d1, nullptr);
    k_ice_hi<<<grid, block, 0, s>>>(a, a.d1, a.d0, a.w);
    k_ice_ratio<<<grid, block, 0, s>>>(a);
    k_ice_apply<<<grid, block, 0, s>>>(a);
    return check_launch("k_ice");
}

}  // namespace fct
