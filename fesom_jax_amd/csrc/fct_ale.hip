// Vertical velocity from continuity, compute_w (fesom_jax/ale.py:309-365): the
// per-level antisymmetric edge divergence summed into each node, the bottom-to-top
// cumulative sum over the column, and the division by the interface area.
//
// Mapping: one 64-lane wavefront per node, lane = level (two slots for nl <= 128). The
// divergence sums the node's (edge, side) list in increasing 2e+side order from +0.0,
// the order of the reference's flattened segment sum (ops.py:56-81). The cumulative
// sum is sequential from the bottom, w[k] = w[k+1] + div[k], the order of the original
// C loop (for nz = nzmax-1 .. nzmin: w[nz] += w[nz+1]); each step reads lane k's
// divergence with readlane, so the running sum is wave-uniform.

#include <hip/hip_runtime.h>

#include "fct_common.h"

namespace fct {
namespace {

template <class T>
using gp = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gp<T> glb(T* p) { return (gp<T>)p; }

__device__ __forceinline__ double read_lane(double v, int lane) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

template <int NS>
__global__ void __launch_bounds__(256) k_compute_w(DevMesh m, const double* uv_,
                                                   const double* helem_, double* w_) {
    const int lane = threadIdx.x & 63;
    const int n = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (n >= m.N) return;
    const int L = m.L;
    const auto uv = glb(uv_);
    const auto helem = glb(helem_);
    const auto w = glb(w_);
    const auto off = glb(m.ecsr_off), lst = glb(m.ecsr);
    const auto etri = glb(reinterpret_cast<const int*>(m.edge_tri));
    const auto elev = glb(reinterpret_cast<const int*>(m.elev));
    const auto cross = glb(reinterpret_cast<const double*>(m.cross));
    const auto area = glb(m.area);
    const int s0 = off[n], s1 = off[n + 1];
    const int nzmin = glb(reinterpret_cast<const int*>(m.nlev))[2 * n];
    const int nzmax = glb(reinterpret_cast<const int*>(m.nlev))[2 * n + 1];   // iface range end

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

    // stage 2: w[k] = w[k+1] + div[k], from the last level up to the surface interface
    // (the divergence is zero at and below nzmax on a consistent mesh, so the sum
    // starts at the bottom interface in practice; starting at the last level keeps
    // the oracle's reversed cumulative sum exactly, signed zeros included)
    double wk[NS];
    for (int s = 0; s < NS; ++s) wk[s] = 0.0;
    double acc = 0.0;
    for (int k = L - 1; k >= nzmin; --k) {
        const int s = k >> 6;
        double d = 0.0;
        for (int t = 0; t < NS; ++t)
            if (t == s) d = read_lane(div[t], k & 63);
        acc = k == L - 1 ? d : acc + d;
        for (int t = 0; t < NS; ++t)
            if (t == s && lane == (k & 63)) wk[t] = acc;
    }

    // stage 3: / area (safe), zero outside the interface range [nzmin, nzmax]
    for (int s = 0; s < NS; ++s) {
        const int k = lane + 64 * s;
        if (k >= L) continue;
        const size_t i = (size_t)n * L + k;
        const double a = area[i];
        const double v = wk[s] / (a > 0.0 ? a : 1.0);
        w[i] = (k >= nzmin && k <= nzmax) ? v : 0.0;
    }
}

}  // namespace

int launch_compute_w(const DevMesh& m, const double* uv, const double* helem, double* w,
                     hipStream_t s) {
    if (m.N == 0) return FCT_OK;
    const int wpb = 4;
    const dim3 grid((m.N + wpb - 1) / wpb), block(64 * wpb);
    if (m.L <= 64)
        k_compute_w<1><<<grid, block, 0, s>>>(m, uv, helem, w);
    else if (m.L <= 128)
        k_compute_w<2><<<grid, block, 0, s>>>(m, uv, helem, w);
    else
        return set_error(FCT_E_UNSUPPORTED, "compute_w supports nl <= 128 (got %d)", m.L);
    return check_launch("k_compute_w");
}

}  // namespace fct
