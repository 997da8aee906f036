// Vertical velocity from continuity, compute_w (fesom_jax/ale.py:309-365): the
// per-level antisymmetric edge divergence summed into each node, the bottom-to-top
// cumulative sum over the column, and the division by the interface area.
//
// Mapping: one 64-lane wavefront per node, lane = level (two slots for nl <= 128). The
// divergence sums the node's (edge, side) list in increasing 2e+side order from +0.0,
// the order of the reference's flattened segment sum (ops.py:56-81). The cumulative
// sum is sequential from the bottom, w[k] = w[k+1] + div[k], the order of the original
// C loop (for nz = nzmax-1 .. nzmin: w[nz] += w[nz+1]), run by one lane per column on
// the columns staged in LDS.
//
// Two forms of the divergence. k_compute_w evaluates each (edge, side) term from the
// triangles' uv and helem, so every edge term is computed at both ends of the edge.
// k_compute_w_vf (meshes with the fused node tables) takes the terms from the FCT's
// edge volume flux, computed once per edge by kf_vflux:
//   term(e, k) = (v1 dx1 - u1 dy1) h1 - (v2 dx2 - u2 dy2) h2 = -vflux(e, k)
// exactly: x - y = -(y - x) and (-x) h = -(x h) in IEEE arithmetic, and a masked
// triangle gives +0 in both. The two can differ only in the sign of a zero, which a
// sum that starts at +0.0 absorbs, so both kernels give the same bits.

#include <hip/hip_runtime.h>

#include "fct_common.h"

namespace fct {
namespace {

template <class T>
using gp = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gp<T> glb(T* p) { return (gp<T>)p; }
// wave-uniform loads of mesh tables through the constant address space (scalar loads)
__device__ __forceinline__ int ldc(const int* p, size_t i) {
    return ((__attribute__((address_space(4))) const int*)p)[i];
}

// stage 2 and 3 of both kernels for node n (valid: n < nod2D), whose divergence is div.
// Blocks are FCT_W_WPB wavefronts, one node each; every thread reaches the barriers.
constexpr int FCT_W_WPB = 4;
template <int NS>
__device__ __forceinline__ void finish_w(const DevMesh& m, int n, bool valid, int lane,
                                         const double (&div)[NS], double* w_) {
    __shared__ double s_col[FCT_W_WPB][128];
    const int L = m.L, wv = (int)(threadIdx.x >> 6);
    const auto nlev = glb(reinterpret_cast<const int*>(m.nlev));
    // stage 2: w[k] = w[k+1] + div[k], from the last level up to the surface interface
    // (the divergence is zero at and below nzmax on a consistent mesh, so the sum
    // starts at the bottom interface in practice; starting at the last level keeps
    // the oracle's reversed cumulative sum exactly, signed zeros included). The columns
    // go through LDS so that one lane per column runs the serial sum.
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int k = lane + 64 * s;
        if (k < L) s_col[wv][k] = div[s];
    }
    __syncthreads();
    if (threadIdx.x < FCT_W_WPB) {
        const int c = (int)threadIdx.x;
        const int nn = (int)blockIdx.x * FCT_W_WPB + c;
        if (nn < m.N) {
            const int zmin = nlev[2 * nn];
            double acc = 0.0;
#pragma unroll 8
            for (int k = L - 1; k >= zmin; --k) {
                const double d = s_col[c][k];
                acc = k == L - 1 ? d : acc + d;
                s_col[c][k] = acc;
            }
        }
    }
    __syncthreads();
    if (!valid) return;
    const auto w = glb(w_);
    const auto area = glb(m.area);
    const int nzmin = nlev[2 * n];
    const int nzmax = nlev[2 * n + 1];   // iface range end
    // stage 3: / area (safe), zero outside the interface range [nzmin, nzmax]
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int k = lane + 64 * s;
        if (k >= L) continue;
        const double wk = k >= nzmin ? s_col[wv][k] : 0.0;
        const size_t i = (size_t)n * L + k;
        const double a = area[i];
        const double v = wk / (a > 0.0 ? a : 1.0);
        w[i] = (k >= nzmin && k <= nzmax) ? v : 0.0;
    }
}

// divergence from the edge volume flux (see the top of the file); the node's edge slots
// come from the fused node table, in the same 2e+side order as the CSR list
template <int NS, int DM>
__global__ void __launch_bounds__(256) k_compute_w_vf(DevMesh m, NodeTab tab,
                                                      const int2* ekr_, const double* vflux_,
                                                      double* w_) {
    const int lane = threadIdx.x & 63;
    // the node index is wave-uniform: scalar registers, so the table loads are scalar
    const int n0 = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    const bool valid = n0 < m.N;
    const int n = valid ? n0 : m.N - 1;
    const int L = m.L;
    const auto vflux = glb(vflux_);
    const int* ekr = reinterpret_cast<const int*>(ekr_);
    const int* ent = reinterpret_cast<const int*>(tab.edge);
    const int cnt = ldc(reinterpret_cast<const int*>(tab.hdr), 4 * (size_t)n + 2) & 0xff;
    int code[DM], lo[DM], hi[DM];
#pragma unroll
    for (int q = 0; q < DM; ++q) {
        code[q] = ldc(ent, 4 * ((size_t)n * DM + q));
        lo[q] = ldc(ekr, 2 * (size_t)(code[q] >> 1));
        hi[q] = ldc(ekr, 2 * (size_t)(code[q] >> 1) + 1);
    }
    double div[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int k = lane + 64 * s;
        const int kc = k < L ? k : L - 1;
        double vf[DM];
#pragma unroll
        for (int q = 0; q < DM; ++q) vf[q] = vflux[(size_t)(code[q] >> 1) * L + kc];
        double d = 0.0;
#pragma unroll
        for (int q = 0; q < DM; ++q) {
            const bool use = q < cnt && k >= lo[q] && k < hi[q];
            d = d + (use ? ((code[q] & 1) ? vf[q] : -vf[q]) : 0.0);
        }
        div[s] = d;
    }
    finish_w<NS>(m, n, valid, lane, div, w_);
}

template <int NS>
__global__ void __launch_bounds__(256) k_compute_w(DevMesh m, const double* uv_,
                                                   const double* helem_, double* w_) {
    const int lane = threadIdx.x & 63;
    const int n0 = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const bool valid = n0 < m.N;
    const int n = valid ? n0 : m.N - 1;
    const int L = m.L;
    const auto uv = glb(uv_);
    const auto helem = glb(helem_);
    const auto off = glb(m.ecsr_off), lst = glb(m.ecsr);
    const auto etri = glb(reinterpret_cast<const int*>(m.edge_tri));
    const auto elev = glb(reinterpret_cast<const int*>(m.elev));
    const auto cross = glb(reinterpret_cast<const double*>(m.cross));
    const int s0 = off[n], s1 = off[n + 1];

    // stage 1: div[k] = sum over the node's (edge, side) of +-c[e][k]
    double div[NS];
    for (int s = 0; s < NS; ++s) div[s] = 0.0;
    for (int i = s0; i < s1; ++i) {
        const int code = lst[i];
        const int e = code >> 1;
        const int el1 = etri[2 * e], el2 = etri[2 * e + 1];
        const double dx1 = cross[4 * (size_t)e], dy1 = cross[4 * (size_t)e + 1];
        const double dx2 = cross[4 * (size_t)e + 2], dy2 = cross[4 * (size_t)e + 3];
        const int a1 = el1 >= 0 ? el1 : 0, a2 = el2 >= 0 ? el2 : 0;
        const int lo1 = elev[2 * a1], hi1 = elev[2 * a1 + 1];
        const int lo2 = elev[2 * a2], hi2 = elev[2 * a2 + 1];
        for (int s = 0; s < NS; ++s) {
            const int k = lane + 64 * s;
            if (k >= L) continue;
            const size_t i1 = (size_t)a1 * L + k, i2 = (size_t)a2 * L + k;
            const bool m1 = el1 >= 0 && k >= lo1 && k < hi1;
            const bool m2 = el2 >= 0 && k >= lo2 && k < hi2;
            // loads are unconditional (the indices are in bounds) and masked afterwards,
            // so all of them are in flight at once
            const double u1 = uv[2 * i1], v1 = uv[2 * i1 + 1], h1 = helem[i1];
            const double u2 = uv[2 * i2], v2 = uv[2 * i2 + 1], h2 = helem[i2];
            const double t1 = m1 ? (v1 * dx1 - u1 * dy1) * h1 : 0.0;
            const double t2 = m2 ? (v2 * dx2 - u2 * dy2) * h2 : 0.0;
            const double c = t1 + -t2;              // sign * where(mask, term, 0), summed
            div[s] = div[s] + ((code & 1) ? -c : c);
        }
    }

    finish_w<NS>(m, n, valid, lane, div, w_);
}

}  // namespace

int launch_compute_w(const DevMesh& m, const double* uv, const double* helem, double* w,
                     hipStream_t s) {
    if (m.N == 0) return FCT_OK;
    const int wpb = FCT_W_WPB;
    const dim3 grid((m.N + wpb - 1) / wpb), block(64 * wpb);
    if (m.L <= 64)
        k_compute_w<1><<<grid, block, 0, s>>>(m, uv, helem, w);
    else if (m.L <= 128)
        k_compute_w<2><<<grid, block, 0, s>>>(m, uv, helem, w);
    else
        return set_error(FCT_E_UNSUPPORTED, "compute_w supports nl <= 128 (got %d)", m.L);
    return check_launch("k_compute_w");
}

int launch_compute_w_vf(const DevMesh& m, const NodeTab& tab, const int2* ekr,
                        const double* vflux, double* w, hipStream_t s) {
    if (m.N == 0) return FCT_OK;
    const int wpb = FCT_W_WPB;
    const dim3 grid((m.N + wpb - 1) / wpb), block(64 * wpb);
    if (m.L > 128) return set_error(FCT_E_UNSUPPORTED, "compute_w supports nl <= 128 (got %d)", m.L);
    if (tab.dm == 6 && m.L <= 64) k_compute_w_vf<1, 6><<<grid, block, 0, s>>>(m, tab, ekr, vflux, w);
    else if (tab.dm == 6) k_compute_w_vf<2, 6><<<grid, block, 0, s>>>(m, tab, ekr, vflux, w);
    else if (tab.dm == 8 && m.L <= 64) k_compute_w_vf<1, 8><<<grid, block, 0, s>>>(m, tab, ekr, vflux, w);
    else if (tab.dm == 8) k_compute_w_vf<2, 8><<<grid, block, 0, s>>>(m, tab, ekr, vflux, w);
    else return set_error(FCT_E_UNSUPPORTED, "compute_w from vflux needs the fused node table");
    return check_launch("k_compute_w_vf");
}

}  // namespace fct
