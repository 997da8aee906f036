// Sea-ice FCT advection of the three ice tracers (fct_solve, fesom_jax/ice_adv.py:37-194):
// Taylor-Galerkin right-hand side, consistent-mass low- and high-order solves, element
// antidiffusive fluxes and the Zalesak limiter, on the surface nodes.
//
// Mapping: one thread per node and all three tracers per thread (the problem is 2-D;
// the tracers share every geometry load). Each node gathers over its elements in
// increasing element id (the node -> element CSR), which is the order of the
// reference's scatter-add over (element, vertex); sums over an element's three
// vertices run ((x0 + x1) + x2). Five launches, each a gather pass over the elements
// around every node: right-hand side + low-order solve, two high-order corrections (the
// second also stores each node's gamma * X + d, the vertex value of the antidiffusive
// fluxes), fluxes + bounds + ratios, limited fluxes + update.

#include <hip/hip_runtime.h>

#include "fct_common.h"

namespace fct {
namespace {

template <class T>
using gp = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gp<T> glb(T* p) { return (gp<T>)p; }

constexpr double ICE_FLUX_EPS = 1.0e-12;   // ice_adv.py:32
constexpr int ICE_K = 3;                   // a_ice, m_ice, m_snow

struct IceArgs {
    int N;
    int L;
    double dt, diff, gamma;
    const int* ncsr_off;
    const int* ncsr;
    const int* en;          // [E*3]
    const double* gs;       // [E*6]
    const double* vol;      // [E]
    const double* sq;       // [E] sqrt(vol / 2e8)
    const int* elev;        // [E*2] (ulevels-1, nlevels-1)
    const int* nlev;        // [N*2] (ulevels_nod2D-1, nlevels_nod2D-1)
    const double* area;     // [N*L]; the surface column is area[n*L]
    const double* u;
    const double* v;
    const double* X[ICE_K];     // tracers in
    double* out[ICE_K];         // tracers out
    double* rhs;            // [3][N]
    double* xl;             // [3][N] low-order solution
    double* d0;             // [3][N] high-order increment (ping)
    double* d1;             // [3][N] (pong)
    double* pp;             // [3][N] ratios
    double* pm;
    double* inv;            // [N] 1 / surface area (k_ice_rhs), as inv_area computes it
    double* w;              // [3][N] gamma * X + d with the final d (second k_ice_hi)
};

__device__ __forceinline__ double inv_area(const IceArgs& a, int n) {
    const double ar = glb(a.area)[(size_t)n * a.L];
    return 1.0 / (ar > 0 ? ar : 1.0);
}
__device__ __forceinline__ bool node_open(const IceArgs& a, int n) {
    return glb(a.nlev)[2 * n] == 0;            // ulevels_nod2D <= 1
}
__device__ __forceinline__ bool elem_open(const IceArgs& a, int e) {
    return glb(a.elev)[2 * e] == 0;            // ulevels <= 1
}
__device__ __forceinline__ int slot_of(int n, int v0, int v1) { return v0 == n ? 0 : (v1 == n ? 1 : 2); }
__device__ __forceinline__ double sum3(double x0, double x1, double x2) { return (x0 + x1) + x2; }
__device__ __forceinline__ double min1(double x) { return x > 1.0 ? 1.0 : x; }   // jnp.minimum(x, 1)

// (mm X)[n]'s term of element e: vol/12 (X[n] + sum of the vertices)
__device__ __forceinline__ double mm_term(double vol, double xq, double x0, double x1, double x2) {
    return vol / 12.0 * (xq + sum3(x0, x1, x2));
}

// pass 1: right-hand side (ice_adv.py:47-73), low-order solution (:76-87) and the first
// high-order approximation rhs / area (:101)
__global__ void __launch_bounds__(256) k_ice_rhs(IceArgs a) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.N) return;
    const auto en = glb(a.en);
    const auto gs = glb(a.gs);
    const auto U = glb(a.u), V = glb(a.v);
    double rhs[ICE_K] = {0.0, 0.0, 0.0}, mm[ICE_K] = {0.0, 0.0, 0.0}, xn[ICE_K];
    for (int t = 0; t < ICE_K; ++t) xn[t] = glb(a.X[t])[n];
    const int s1 = glb(a.ncsr_off)[n + 1];
    for (int s = glb(a.ncsr_off)[n]; s < s1; ++s) {
        const int e = glb(a.ncsr)[s];
        const int v0 = en[3 * e], v1 = en[3 * e + 1], v2 = en[3 * e + 2];
        const int q = slot_of(n, v0, v1);
        const bool on = elem_open(a, e);
        const double vol = glb(a.vol)[e];
        double dx[3], dy[3];
        for (int i = 0; i < 3; ++i) {
            dx[i] = gs[6 * e + i];
            dy[i] = gs[6 * e + 3 + i];
        }
        const double Uv[3] = {U[v0], U[v1], U[v2]}, Vv[3] = {V[v0], V[v1], V[v2]};
        const double um = sum3(Uv[0], Uv[1], Uv[2]), vm = sum3(Vv[0], Vv[1], Vv[2]);
        const double diff = a.diff * glb(a.sq)[e];
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
            const auto X = glb(a.X[t]);
            const double x0 = X[v0], x1 = X[v1], x2 = X[v2];
            const double c = (ent[0] * x0 + ent[1] * x1) + ent[2] * x2;
            const double xq = q == 0 ? x0 : (q == 1 ? x1 : x2);
            rhs[t] = rhs[t] + (on ? c : 0.0);
            mm[t] = mm[t] + (on ? mm_term(vol, xq, x0, x1, x2) : 0.0);
        }
    }
    const double inv = inv_area(a, n);
    glb(a.inv)[n] = inv;
    const bool nc = node_open(a, n);
    for (int t = 0; t < ICE_K; ++t) {
        const size_t i = (size_t)t * a.N + n;
        glb(a.rhs)[i] = rhs[t];
        glb(a.xl)[i] = nc ? (rhs[t] + a.gamma * mm[t]) * inv + (1.0 - a.gamma) * xn[t] : xn[t];
        glb(a.d0)[i] = nc ? rhs[t] * inv : 0.0;
    }
}

// pass 2 (twice): d <- d + (rhs - mm d) / area (ice_adv.py:102-104); the second pass also
// stores gamma * X + d (wout), the vertex value of the antidiffusive fluxes
__global__ void __launch_bounds__(256) k_ice_hi(IceArgs a, const double* din, double* dout,
                                                double* wout) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.N) return;
    const auto en = glb(a.en);
    const auto D = glb(din);
    double mm[ICE_K] = {0.0, 0.0, 0.0};
    const int s1 = glb(a.ncsr_off)[n + 1];
    for (int s = glb(a.ncsr_off)[n]; s < s1; ++s) {
        const int e = glb(a.ncsr)[s];
        const int v0 = en[3 * e], v1 = en[3 * e + 1], v2 = en[3 * e + 2];
        const int q = slot_of(n, v0, v1);
        const bool on = elem_open(a, e);
        const double vol = glb(a.vol)[e];
        for (int t = 0; t < ICE_K; ++t) {
            const size_t o = (size_t)t * a.N;
            const double x0 = D[o + v0], x1 = D[o + v1], x2 = D[o + v2];
            const double xq = q == 0 ? x0 : (q == 1 ? x1 : x2);
            mm[t] = mm[t] + (on ? mm_term(vol, xq, x0, x1, x2) : 0.0);
        }
    }
    const double inv = inv_area(a, n);
    const bool nc = node_open(a, n);
    for (int t = 0; t < ICE_K; ++t) {
        const size_t i = (size_t)t * a.N + n;
        const double d = D[i];
        const double dn = nc ? d + (glb(a.rhs)[i] - mm[t]) * inv : d;
        glb(dout)[i] = dn;
        if (wout) glb(wout)[i] = a.gamma * glb(a.X[t])[n] + dn;
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
    const auto en = glb(a.en);
    const auto W = glb(a.w);
    const auto XL = glb(a.xl);
    const double inv = inv_area(a, n);
    double fp[ICE_K] = {0.0, 0.0, 0.0}, fm[ICE_K] = {0.0, 0.0, 0.0};
    double hi[ICE_K], lo[ICE_K];
    for (int t = 0; t < ICE_K; ++t) { hi[t] = -INFINITY; lo[t] = INFINITY; }
    const int s1 = glb(a.ncsr_off)[n + 1];
    for (int s = glb(a.ncsr_off)[n]; s < s1; ++s) {
        const int e = glb(a.ncsr)[s];
        const int vv[3] = {en[3 * e], en[3 * e + 1], en[3 * e + 2]};
        const int q = slot_of(n, vv[0], vv[1]);
        const bool on = elem_open(a, e);
        const double vol = glb(a.vol)[e];
        for (int t = 0; t < ICE_K; ++t) {
            const size_t o = (size_t)t * a.N;
            const auto X = glb(a.X[t]);
            double w[3], eh = -INFINITY, el = INFINITY;
            for (int j = 0; j < 3; ++j) {
                const double x = X[vv[j]], l = XL[o + vv[j]];
                w[j] = W[o + vv[j]];
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
        const double xl = XL[i];
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
    const auto en = glb(a.en);
    const auto W = glb(a.w);
    const auto PP = glb(a.pp), PM = glb(a.pm);
    double acc[ICE_K] = {0.0, 0.0, 0.0};
    const int s1 = glb(a.ncsr_off)[n + 1];
    for (int s = glb(a.ncsr_off)[n]; s < s1; ++s) {
        const int e = glb(a.ncsr)[s];
        const int vv[3] = {en[3 * e], en[3 * e + 1], en[3 * e + 2]};
        const int q = slot_of(n, vv[0], vv[1]);
        const bool on = elem_open(a, e);
        const double vol = glb(a.vol)[e];
        double invj[3];
        for (int j = 0; j < 3; ++j) invj[j] = glb(a.inv)[vv[j]];
        for (int t = 0; t < ICE_K; ++t) {
            const size_t o = (size_t)t * a.N;
            double w[3];
            for (int j = 0; j < 3; ++j) w[j] = W[o + vv[j]];
            // the element's factor takes each vertex's ratio by the sign of that vertex's
            // flux (0 when the element is closed); only the node's own flux is divided
            double ae = INFINITY, gq = 0.0;
            for (int j = 0; j < 3; ++j) {
                const double g = ice_flux_num(w, j, vol, invj[j]);
                gq = j == q ? g : gq;
                ae = fmin(ae, !on || ice_flux_nonneg(g) ? PP[o + vv[j]] : PM[o + vv[j]]);
            }
            acc[t] = acc[t] + (on ? (gq / 12.0) * ae : 0.0);
        }
    }
    for (int t = 0; t < ICE_K; ++t) {
        const size_t i = (size_t)t * a.N + n;
        glb(a.out[t])[n] = glb(a.xl)[i] + acc[t];
    }
}

}  // namespace

size_t ice_workspace_doubles(int N) { return (size_t)22 * N; }

int launch_ice_fct(const DevMesh& m, const IceCall& c, double* work, const double* sq,
                   hipStream_t s) {
    if (m.N == 0) return FCT_OK;
    IceArgs a{};
    a.N = m.N;
    a.L = m.L;
    a.dt = c.dt;
    a.diff = c.diff;
    a.gamma = c.gamma;
    a.ncsr_off = m.ncsr_off;
    a.ncsr = m.ncsr;
    a.en = m.en;
    a.gs = m.gs;
    a.vol = m.earea;
    a.sq = sq;
    a.elev = reinterpret_cast<const int*>(m.elev);
    a.nlev = reinterpret_cast<const int*>(m.nlev);
    a.area = m.area;
    a.u = c.u;
    a.v = c.v;
    for (int t = 0; t < ICE_K; ++t) {
        a.X[t] = c.X[t];
        a.out[t] = c.out[t];
    }
    const size_t N3 = (size_t)3 * m.N;
    a.rhs = work;
    a.xl = work + N3;
    a.d0 = work + 2 * N3;
    a.d1 = work + 3 * N3;
    a.pp = work + 4 * N3;
    a.pm = work + 5 * N3;
    a.inv = work + 6 * N3;
    a.w = work + 6 * N3 + m.N;
    const dim3 grid((m.N + 255) / 256), block(256);
    k_ice_rhs<<<grid, block, 0, s>>>(a);
    k_ice_hi<<<grid, block, 0, s>>>(a, a.d0, a.d1, nullptr);
    k_ice_hi<<<grid, block, 0, s>>>(a, a.d1, a.d0, a.w);
    k_ice_ratio<<<grid, block, 0, s>>>(a);
    k_ice_apply<<<grid, block, 0, s>>>(a);
    return check_launch("k_ice");
}

}  // namespace fct
