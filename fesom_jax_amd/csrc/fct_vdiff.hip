// Implicit vertical tracer diffusion: one tridiagonal solve per node column, shared by
// up to two tracers (impl_vert_diff / impl_vert_diff_one, tracer_diff.py:68-141), with
// the optional lower bound on wet layers that the step applies to salinity right after
// it (step.py:456).
//
// Mapping: a workgroup is one 64-lane wavefront; lane = node column. The Thomas
// recurrence (ops.py:121-168) runs down each lane's column in the reference's operation
// order, over the rows [0, Lw) where Lw is the deepest bottom in the wavefront (rows
// further down are dry in every lane and solve to +0 in the reference too). The inputs
// are node-major ([node][level]); they are staged through LDS in chunks of 8 levels, with
// lanes spread over (column, level) for the global loads so that eight lanes read one
// column's 64 contiguous bytes. The forward sweep keeps cp and dp in a level-major
// scratch slot of the workgroup ([level][64 lanes], coalesced); the grid is sized to the
// workgroups the GPU holds at once, each looping over column blocks, so the scratch is
// bounded by that residency, not by the mesh.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "fct_common.h"

namespace fct {
namespace {

constexpr int VD_COLS = 64;      // columns per workgroup (one wavefront)
constexpr int VD_KC = 8;         // levels per staged chunk
constexpr int VD_ROW = 66;       // LDS row stride in doubles (spreads the transposed writes)
constexpr int VD_SLOT = VD_KC * VD_ROW;
#ifndef VD_OCC
#define VD_OCC 1                 // wavefronts per SIMD the register budget is sized for
#endif

struct VdArgs {
    int N, L, nblk;
    double dt;
    const int2* nlev;
    const double* area;
    const double* av;
    const double* Zp;             // static layer depths [L]: Z[k], Z[nl-2] at k = nl-1
    const double* Kv;
    const double* hnn;
    const double* Z3d;
    const double* T[2];
    const double* bc[2];
    const double* sw[2];
    double* Tn[2];
    int floor_j;                  // tracer with a lower bound on wet layers, -1 = none
    double floor_v;
    double* scratch;              // per workgroup: (1 + K) * L * 64 doubles
};

// Global-address-space views of the field pointers: loaded from a struct, they are
// generic pointers, which could alias the LDS staging buffer and would keep every
// global load behind the preceding LDS store.
template <class T>
using gp = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gp<T> glb(T* p) { return (gp<T>)p; }

// Barrier that orders only the LDS staging buffer: the workgroup's global stores (the
// scratch and the outputs) need no completion here, each lane re-reads only its own.
__device__ __forceinline__ void lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// LDS slots: shifted fields hold level k+1 for row k, the others level k.
enum { S_KV = 0, S_AREA, S_AV, S_HNN, S_T0, S_T1 };

template <int K, bool Z3D, bool SW>
struct VdLayout {
    static constexpr int kZ = S_T0 + K;
    static constexpr int kSW = kZ + (Z3D ? 1 : 0);
    static constexpr int count = kSW + (SW ? K : 0);
};

// Load / store at a 32-bit byte offset from a wave-uniform base: the base stays in
// scalar registers and each access costs one 32-bit vector register for its offset.
template <class T>
__device__ __forceinline__ T ld_at(gp<const T> base, uint32_t off) {
    return *(gp<const T>)((gp<const char>)base + off);
}
template <class T>
__device__ __forceinline__ void st_at(gp<T> base, uint32_t off, T v) {
    *(gp<T>)((gp<char>)base + off) = v;
}

// Hides an offset's derivation from loop strength reduction, which would otherwise
// keep one 64-bit pointer per field and column group alive across the whole sweep.
__device__ __forceinline__ uint32_t opaque(uint32_t x) {
    asm volatile("" : "+v"(x));
    return x;
}

// The fields of one column block (columns n0 .. n0+63), as wave-uniform bases.
template <int K>
struct VdBlock {
    gp<const double> av, hnn, Kv, area, Z3d;
    gp<const double> T[K], sw[K];
    gp<double> Tn[K];
    __device__ VdBlock(const VdArgs& a, int n0) {
        const size_t o = (size_t)n0 * a.L;
        av = glb(a.av) + o;
        hnn = glb(a.hnn) + o;
        Kv = glb(a.Kv) + o;
        area = glb(a.area) + o;
        Z3d = glb(a.Z3d) + o;
        for (int j = 0; j < K; ++j) {
            T[j] = glb(a.T[j]) + o;
            sw[j] = glb(a.sw[j]) + o;
            Tn[j] = glb(a.Tn[j]) + o;
        }
    }
};

constexpr int VD_NI = VD_COLS / VD_KC;   // column groups per chunk (lanes = 8 levels x 8 columns)

// One chunk of the forward sweep's inputs, in registers between the loads and the
// transposed LDS writes: lane (q, r) holds level k0+q (k0+q+1 for the shifted fields) of
// columns r, r+8, ... Every load is unconditional from an in-bounds offset and masked
// when written: a guarded load compiles to a branch, which keeps the loads after it from
// issuing before it returns. A block's offsets stay below 64 * nl * 8 bytes.
template <int K, bool Z3D, bool SW>
struct VdChunk {
    double av[VD_NI], hnn[VD_NI], kv[VD_NI], ar[VD_NI], t[K][VD_NI], sw[K][VD_NI], z[VD_NI];
    double zq;                    // static depth Zp[k0+q+1]

    __device__ void load(const VdBlock<K>& b, const VdArgs& a, int n0, int k0, int q, int r) {
        const int L = a.L;
#pragma unroll
        for (int i = 0; i < VD_NI; ++i) {
            const int c = VD_KC * i + r;
            const bool okc = n0 + c < a.N;
            const int ku = k0 + q, ks = k0 + q + 1;
            const uint32_t iu = opaque((okc && ku < L) ? (uint32_t)(c * L + ku) * 8u : 0u);
            const uint32_t is = opaque((okc && ks < L) ? (uint32_t)(c * L + ks) * 8u : 0u);
            av[i] = ld_at(b.av, iu);
            hnn[i] = ld_at(b.hnn, iu);
            kv[i] = ld_at(b.Kv, is);
            ar[i] = ld_at(b.area, is);
            for (int j = 0; j < K; ++j) {
                t[j][i] = ld_at(b.T[j], is);
                sw[j][i] = 0.0;
                if (SW && a.sw[j]) sw[j][i] = ld_at(b.sw[j], is);
            }
            z[i] = Z3D ? ld_at(b.Z3d, is) : 0.0;
        }
        zq = Z3D ? 0.0 : glb(a.Zp)[min(k0 + q + 1, L - 1)];
    }

    __device__ void put(double* sh, double* szs, const VdArgs& a, int n0, int k0, int q,
                        int r) const {
        using Lay = VdLayout<K, Z3D, SW>;
#pragma unroll
        for (int i = 0; i < VD_NI; ++i) {
            const int c = VD_KC * i + r;
            const bool okc = n0 + c < a.N;
            const bool oku = okc && k0 + q < a.L, oks = okc && k0 + q + 1 < a.L;
            double* d = sh + q * VD_ROW + c;
            d[S_AV * VD_SLOT] = oku ? av[i] : 0.0;
            d[S_HNN * VD_SLOT] = oku ? hnn[i] : 0.0;
            d[S_KV * VD_SLOT] = oks ? kv[i] : 0.0;
            d[S_AREA * VD_SLOT] = oks ? ar[i] : 0.0;
            for (int j = 0; j < K; ++j) d[(S_T0 + j) * VD_SLOT] = oks ? t[j][i] : 0.0;
            if (Z3D) d[Lay::kZ * VD_SLOT] = oks ? z[i] : 0.0;
            if (SW)
                for (int j = 0; j < K; ++j) d[(Lay::kSW + j) * VD_SLOT] = oks ? sw[j][i] : 0.0;
        }
        if (r == 0) szs[q] = (k0 + q + 1 < a.L) ? zq : 0.0;
    }
};

// One chunk of the back substitution's inputs: cp and dp of rows k0..k0+7 of the
// lane's column (scratch, level-major) and T at (column r + 8i, level k0+q) for the
// output, loaded together.
template <int K>
struct VdBack {
    double cp[VD_KC], dp[K][VD_KC], tv[K][VD_NI];

    __device__ void load(const VdBlock<K>& b, gp<double> scr, const VdArgs& a, int n0, int k0,
                         int Lw, int lane, int q, int r) {
        const int L = a.L;
        if (k0 < Lw) {
#pragma unroll
            for (int qq = 0; qq < VD_KC; ++qq) {
                const uint32_t o = opaque((uint32_t)(min(k0 + qq, L - 1) * VD_COLS + lane) * 8u);
                cp[qq] = ld_at<double>(scr, o);
                for (int j = 0; j < K; ++j)
                    dp[j][qq] = ld_at<double>(scr, o + (uint32_t)((1 + j) * L * VD_COLS) * 8u);
            }
        }
#pragma unroll
        for (int i = 0; i < VD_NI; ++i) {
            const int c = VD_KC * i + r, k = k0 + q;
            const uint32_t g = opaque((n0 + c < a.N && k < L) ? (uint32_t)(c * L + k) * 8u : 0u);
            for (int j = 0; j < K; ++j) tv[j][i] = ld_at(b.T[j], g);
        }
    }
};

template <int K, bool Z3D, bool SW>
__global__ void __launch_bounds__(VD_COLS) __attribute__((amdgpu_waves_per_eu(VD_OCC))) k_vdiff(VdArgs a) {
    using Lay = VdLayout<K, Z3D, SW>;
    extern __shared__ double sh[];
    __shared__ int2 s_nlev[VD_COLS];
    __shared__ double s_zs[VD_KC];
    const int lane = threadIdx.x;
    const int q = lane & (VD_KC - 1), r = lane >> 3;
    const int L = a.L;
    const gp<double> scr = glb(a.scratch + (size_t)blockIdx.x * (1 + K) * L * VD_COLS);
    const gp<const int> g_nlev = glb(reinterpret_cast<const int*>(a.nlev));

    for (int blk = blockIdx.x; blk < a.nblk; blk += gridDim.x) {
        const int n0 = blk * VD_COLS;
        const VdBlock<K> b(a, n0);
        const int n = n0 + lane;
        const bool own = n < a.N;
        const int nc = own ? n : 0;
        const int2 nr = own ? make_int2(g_nlev[2 * n], g_nlev[2 * n + 1]) : make_int2(0, 0);
        lds_sync();              // the previous block's LDS readers are done
        s_nlev[lane] = nr;
        int Lw = nr.y;
        for (int o = 32; o > 0; o >>= 1) Lw = max(Lw, __shfl_xor(Lw, o));
        const uint32_t row = own ? (uint32_t)(lane * L) * 8u : 0u;

        // level-0 values of the shifted fields, carried into row 0
        double Tc[K], Tp[K], swc[K], bcv[K];
        for (int j = 0; j < K; ++j) {
            const double t = ld_at(b.T[j], row);
            Tc[j] = own ? t : 0.0;
            Tp[j] = Tc[j];
            double w = 0.0;
            if (SW && a.sw[j]) w = ld_at(b.sw[j], row);
            swc[j] = own ? w : 0.0;
            double bb = 0.0;
            if (a.bc[j]) bb = glb(a.bc[j])[nc];
            bcv[j] = own ? bb : 0.0;
        }
        const double kv0 = ld_at(b.Kv, row), ar0 = ld_at(b.area, row);
        double Kvc = own ? kv0 : 0.0;
        double arc = own ? ar0 : 0.0;
        double Zc = 0.0;
        if (Z3D) {
            const double z0 = ld_at(b.Z3d, row);
            Zc = own ? z0 : 0.0;
        } else {
            Zc = glb(a.Zp)[0];
        }
        double Zprev = Zc;
        double cp_prev = 0.0, dp_prev[K];
        for (int j = 0; j < K; ++j) dp_prev[j] = 0.0;

        // forward elimination, chunk k0+8's loads in flight while chunk k0 is computed
        VdChunk<K, Z3D, SW> st;
        if (Lw > 0) st.load(b, a, n0, 0, q, r);
        for (int k0 = 0; k0 < Lw; k0 += VD_KC) {
            lds_sync();
            st.put(sh, s_zs, a, n0, k0, q, r);
            lds_sync();
            if (k0 + VD_KC < Lw) st.load(b, a, n0, k0 + VD_KC, q, r);
            const int kend = min(VD_KC, Lw - k0);
#pragma unroll 1
            for (int qq = 0; qq < kend; ++qq) {
                const int k = k0 + qq;
                const double* s = sh + qq * VD_ROW + lane;
                const double av = s[S_AV * VD_SLOT];
                const double h = s[S_HNN * VD_SLOT];
                const double Kvn = s[S_KV * VD_SLOT];
                const double arn = s[S_AREA * VD_SLOT];
                const double Zn = Z3D ? s[Lay::kZ * VD_SLOT] : s_zs[qq];
                double Tn[K], swn[K];
                for (int j = 0; j < K; ++j) {
                    Tn[j] = s[(S_T0 + j) * VD_SLOT];
                    swn[j] = SW ? s[(Lay::kSW + j) * VD_SLOT] : 0.0;
                }
                // coefficients (tracer_diff.py:84-112)
                const double dZu = nz_safe(Zprev - Zc);
                const double dZd = nz_safe(Zc - Zn);
                const double sav = av > 0.0 ? av : 1.0;
                const double af = -Kvc * a.dt / dZu * (arc / sav);
                const double arn_sav = arn / sav;
                const double cf = -Kvn * a.dt / dZd * arn_sav;
                const bool valid = in_layer(nr, k), surf = k == nr.x, bot = k == nr.y - 1;
                double ca = surf ? 0.0 : af;
                double cc = bot ? 0.0 : cf;
                const double cb = valid ? -ca - cc + h : 1.0;
                ca = valid ? ca : 0.0;
                cc = valid ? cc : 0.0;
                // forward elimination (ops.py:156-162)
                const double mm = cb - ca * cp_prev;
                const double cpk = cc / mm;
                const uint32_t o = (uint32_t)(k * VD_COLS + lane) * 8u;
                st_at<double>(scr, o, cpk);
                cp_prev = cpk;
                for (int j = 0; j < K; ++j) {
                    // right-hand side (tracer_diff.py:113-125)
                    double rhs = valid ? ca * (Tc[j] - Tp[j]) + cc * (Tc[j] - Tn[j]) : 0.0;
                    if (a.bc[j]) rhs = rhs + ((surf && valid) ? bcv[j] : 0.0);
                    if (SW && a.sw[j])
                        rhs = rhs + (valid ? (swc[j] - swn[j] * arn_sav) * a.dt : 0.0);
                    const double dpk = (rhs - ca * dp_prev[j]) / mm;
                    st_at<double>(scr, o + (uint32_t)((1 + j) * L * VD_COLS) * 8u, dpk);
                    dp_prev[j] = dpk;
                    Tp[j] = Tc[j];
                    Tc[j] = Tn[j];
                    swc[j] = swn[j];
                }
                Kvc = Kvn;
                arc = arn;
                Zprev = Zc;
                Zc = Zn;
            }
        }

        // back substitution (ops.py:164-167) chunk by chunk from the deepest, then the
        // coalesced write of T + increment on wet layers (T + 0 elsewhere); chunk ci-1's
        // loads are in flight while chunk ci is finished
        double xn[K];
        for (int j = 0; j < K; ++j) xn[j] = 0.0;
        const int nchunk = (L + VD_KC - 1) / VD_KC;
        VdBack<K> bk;
        bk.load(b, scr, a, n0, (nchunk - 1) * VD_KC, Lw, lane, q, r);
        for (int ci = nchunk - 1; ci >= 0; --ci) {
            const int k0 = ci * VD_KC;
            const VdBack<K> cur = bk;
            if (ci > 0) bk.load(b, scr, a, n0, k0 - VD_KC, Lw, lane, q, r);
            lds_sync();
#pragma unroll
            for (int qq = VD_KC - 1; qq >= 0; --qq) {
                const bool live = k0 + qq < Lw;
                for (int j = 0; j < K; ++j) {
                    const double x = live ? cur.dp[j][qq] - cur.cp[qq] * xn[j] : 0.0;
                    xn[j] = live ? x : xn[j];
                    sh[j * VD_SLOT + qq * VD_ROW + lane] = x;
                }
            }
            lds_sync();
#pragma unroll
            for (int i = 0; i < VD_NI; ++i) {
                const int c = VD_KC * i + r;
                const int k = k0 + q;
                const bool wet = in_layer(s_nlev[c], k);
                for (int j = 0; j < K; ++j) {
                    const double x = sh[j * VD_SLOT + q * VD_ROW + c];
                    double out = cur.tv[j][i] + (wet ? x : 0.0);
                    if (j == a.floor_j && wet) out = out < a.floor_v ? a.floor_v : out;
                    if (n0 + c < a.N && k < L) st_at<double>(b.Tn[j], (uint32_t)(c * L + k) * 8u, out);
                }
            }
        }
    }
}

template <int K, bool Z3D, bool SW>
size_t vd_lds() {
    return (size_t)std::max(VdLayout<K, Z3D, SW>::count, K) * VD_SLOT * sizeof(double);
}

template <int K, bool Z3D, bool SW>
int vd_resident(int cus, int* out) {
    int nb = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &nb, k_vdiff<K, Z3D, SW>, VD_COLS, vd_lds<K, Z3D, SW>());
    if (e != hipSuccess)
        return set_error(FCT_E_HIP, "occupancy query failed: %s", hipGetErrorString(e));
    *out = std::max(1, nb) * cus;
    return FCT_OK;
}

template <int K, bool Z3D, bool SW>
int vd_launch(const VdArgs& a, int grid, hipStream_t s) {
    const size_t lds = vd_lds<K, Z3D, SW>();
    k_vdiff<K, Z3D, SW><<<dim3(grid), dim3(VD_COLS), lds, s>>>(a);
    return check_launch("k_vdiff");
}

template <int K>
int vd_dispatch(const VdArgs& a, bool z3d, bool sw, int grid, hipStream_t s) {
    if (z3d) return sw ? vd_launch<K, true, true>(a, grid, s) : vd_launch<K, true, false>(a, grid, s);
    return sw ? vd_launch<K, false, true>(a, grid, s) : vd_launch<K, false, false>(a, grid, s);
}

template <int K>
int vd_resident_k(bool z3d, bool sw, int cus, int* out) {
    if (z3d) return sw ? vd_resident<K, true, true>(cus, out) : vd_resident<K, true, false>(cus, out);
    return sw ? vd_resident<K, false, true>(cus, out) : vd_resident<K, false, false>(cus, out);
}

}  // namespace

int vdiff_max_grid(int cus, int* out) {
    int best = 0;
    for (int K = 1; K <= 2; ++K)
        for (int z = 0; z < 2; ++z)
            for (int w = 0; w < 2; ++w) {
                int g = 0;
                const int rc = K == 1 ? vd_resident_k<1>(z, w, cus, &g)
                                      : vd_resident_k<2>(z, w, cus, &g);
                if (rc) return rc;
                best = std::max(best, g);
            }
    *out = best;
    return FCT_OK;
}

size_t vdiff_scratch_doubles(int L, int grid) { return (size_t)3 * L * VD_COLS * grid; }

int launch_vdiff(const DevMesh& m, const VdiffCall& c, int max_grid, double* scratch,
                 hipStream_t s) {
    VdArgs a{};
    a.N = m.N;
    a.L = m.L;
    a.nblk = (m.N + VD_COLS - 1) / VD_COLS;
    a.dt = c.dt;
    a.nlev = m.nlev;
    a.area = m.area;
    a.av = m.areasvol;
    a.Zp = m.Zs;                    // the first L entries are Zp
    a.Kv = c.Kv;
    a.hnn = c.hnn;
    a.Z3d = c.Z3d;
    a.floor_j = c.floor_j;
    a.floor_v = c.floor_v;
    a.scratch = scratch;
    bool sw = false;
    for (int j = 0; j < c.K; ++j) {
        a.T[j] = c.T[j];
        a.Tn[j] = c.Tn[j];
        a.bc[j] = c.bc[j];
        a.sw[j] = c.sw[j];
        sw = sw || c.sw[j] != nullptr;
    }
    if (a.nblk == 0) return FCT_OK;
    const int grid = std::min(a.nblk, max_grid);
    return c.K == 1 ? vd_dispatch<1>(a, c.Z3d != nullptr, sw, grid, s)
                    : vd_dispatch<2>(a, c.Z3d != nullptr, sw, grid, s);
}

// The scratch of fct_impl_vert_diff_static: a fixed device array, so that call allocates
// nothing; every GPU the library runs on holds its own copy until the process ends. Host
// code never passes the array's name to a kernel (on the host it is only a stand-in for
// the symbol): it uses the address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) double scratch[FCT_VDIFF_STATIC_BYTES / sizeof(double)];

int vdiff_static_scratch(int device, double** out) {
    static std::mutex mu;
    static std::vector<double*> addr;   // per device ordinal
    std::lock_guard<std::mutex> lock(mu);
    if (device < 0) return set_error(FCT_E_INVALID, "device %d", device);
    if ((int)addr.size() <= device) addr.resize((size_t)device + 1, nullptr);
    if (!addr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_error(FCT_E_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s",
                             device, hipGetErrorString(e));
        addr[device] = static_cast<double*>(ptr);
    }
    *out = addr[device];
    return FCT_OK;
}

}  // namespace fct
