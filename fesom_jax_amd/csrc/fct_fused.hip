// Fused FP64 FCT path for gfx950: two edge passes and three node-column passes.
//
// Mapping: one 64-lane wavefront per task (a node or an edge, times a tracer); lane l of
// slot s holds level k = 64*s + l of the task's column, so every load of a node or edge
// column is one coalesced row segment, and vertical neighbours move between lanes with
// DPP wavefront shifts. For nl > 64 the node passes run as two launches over a node list
// (kl_*): columns whose wet layers fit 64 lanes in a one-slot kernel, which needs fewer
// registers, the others in the two-slot kernel. With two tracers on a 6-slot mesh the
// one-slot launches take one wavefront per node for both tracers (kl_pass*_all), so the
// tracer-independent loads are made once. Task indices are wave-uniform (scalar
// registers); each node reads its stencil (incident edges in the reference's summation
// order with the layer ranges of their triangles, incident triangles with their
// vertices) from a per-node table with scalar loads, and then issues the loads of all
// its neighbours' columns back to back. Only wet lanes are computed; T_new is written
// on every lane (dry lanes: T + 0, as the reference's ale_reconstruct does). The kernels
// are bound by the vector memory path (address and data units, L1), so the design
// removes 64-lane loads wherever it can: tables through scalar loads, k-1/k+1 values by
// lane shifts, the node gradient only where an edge reads it, adf_v handed from pass 2
// to pass 3.
//
//   kf_vflux  (edge) : vflux on the edge's wet levels (edge_vflux, tracer_adv.py:69-93)
//   kf_pass1  (node) : LO and the node-averaged gradient gnode (kl_pass1 over a node list)
//                      (compute_fct_lo :298-313, tracer_gradient_elements :316-326,
//                       _node_avg_grad :329-341)
//   kf_adfh   (edge) : adf_h = MFCT high order - upwind low order, once per edge-level
//                      (:96-105, :196-225, fill_up_dn_grad :344-368)
//   kf_pass2  (node) : cluster bounds, P+/P-, limiter factors (zalesak_limit :389-451),
//                      and adf_v (adv_flux_ver_ho :241-295) stored for pass 3
//   kf_pass3  (node) : limited fluxes, flux2dtracer_fct, ale_reconstruct
//                      (:453-466, :469-484, :142-151)
//
// Operation order and per-node summation order are those of fct_phased.hip (and of the
// reference's jnp expressions). Terms the reference adds as exact zeros (edges or
// elements dry at level k, where the reference multiplies a finite value by 0) are
// skipped, which leaves every sum bit-identical while the inputs are finite on the
// lanes the reference multiplies by zero. The cluster bounds use the fact that the
// vertices of the triangles around a node are the node and the far ends of its edges
// (checked on the host when the plan is built): a maximum over the triangles' vertex
// maxima equals the maximum over that vertex set, whatever the grouping.

#include <cstring>

#include "fct_common.h"

namespace fct {

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

// value of lane l-1 (wave_shr:1) of a 64-bit register, lane 0 keeping `first`, and of
// lane l+1 (wave_shl:1), lane 63 keeping `last` (the DPP move leaves a lane whose source
// is out of range at its old value). Called with all 64 lanes active: a lane whose
// source lane is switched off would keep the old value too. The edge value goes in as
// the old value rather than through a select on the lane index, so the compiler has no
// select to turn into a branch around the move (which would switch off the lanes it
// reads from); the empty asm keeps the move where it is written.
__device__ __forceinline__ double wave_prev(double v, double first) {
    int2 x = __builtin_bit_cast(int2, v);
    const int2 o = __builtin_bit_cast(int2, first);
    x.x = __builtin_amdgcn_update_dpp(o.x, x.x, 0x138, 0xf, 0xf, false);
    x.y = __builtin_amdgcn_update_dpp(o.y, x.y, 0x138, 0xf, 0xf, false);
    asm volatile("" : "+v"(x.x), "+v"(x.y));
    return __builtin_bit_cast(double, x);
}
__device__ __forceinline__ double wave_next(double v, double last) {
    int2 x = __builtin_bit_cast(int2, v);
    const int2 o = __builtin_bit_cast(int2, last);
    x.x = __builtin_amdgcn_update_dpp(o.x, x.x, 0x130, 0xf, 0xf, false);
    x.y = __builtin_amdgcn_update_dpp(o.y, x.y, 0x130, 0xf, 0xf, false);
    asm volatile("" : "+v"(x.x), "+v"(x.y));
    return __builtin_bit_cast(double, x);
}
__device__ __forceinline__ double lane_value(double v, int l) {
    int2 x = __builtin_bit_cast(int2, v);
    x.x = __builtin_amdgcn_readlane(x.x, l);
    x.y = __builtin_amdgcn_readlane(x.y, l);
    return __builtin_bit_cast(double, x);
}

// value at level k-1 of a column held as NS slots; `at0` is used at k = 0
template <int NS>
__device__ __forceinline__ void col_prev(const double (&v)[NS], double (&o)[NS], double at0) {
#pragma unroll
    for (int s = 0; s < NS; ++s)
        o[s] = wave_prev(v[s], (s > 0) ? lane_value(v[s > 0 ? s - 1 : 0], 63) : at0);
}

// value at level k+1; `atend` is used past the last slot lane
template <int NS>
__device__ __forceinline__ void col_next(const double (&v)[NS], double (&o)[NS], double atend) {
#pragma unroll
    for (int s = 0; s < NS; ++s)
        o[s] = wave_next(v[s], (s + 1 < NS) ? lane_value(v[s + 1 < NS ? s + 1 : s], 0) : atend);
}

template <int NS>
__device__ __forceinline__ double col_bcast0(const double (&v)[NS]) { return lane_value(v[0], 0); }

__device__ __forceinline__ double ab2(double told, double t) { return AB_OLD * told + AB_NEW * t; }

__device__ __forceinline__ bool in_bytes(int packed, int shift, int k) {
    const int lo = (packed >> shift) & 0xff, hi = (packed >> (shift + 8)) & 0xff;
    return k >= lo && k < hi;
}

// Block -> tile remap. Workgroups are dealt to the X XCDs round-robin (block b to XCD
// b % X), and node and edge order is spatially local, so giving each XCD runs of C
// consecutive tiles keeps most neighbours of a tile in the same L2. remap = X | C << 8;
// C = 0 gives each XCD one contiguous range of the whole grid. With C > 0 the grid's
// tail beyond the last whole round of X * C blocks keeps hardware order, so the map
// is a bijection on [0, gridDim.x) for any grid size.
__device__ __forceinline__ unsigned block_tile(int remap) {
    const unsigned b = blockIdx.x, G = gridDim.x;
    const unsigned X = (unsigned)remap & 0xffu, C = (unsigned)remap >> 8;
    if (X <= 1) return b;
    const unsigned x = b % X, i = b / X;
    if (C == 0) {
        const unsigned q = G / X, r = G % X;
        return x * q + (x < r ? x : r) + i;
    }
    const unsigned span = X * C;
    if (b >= G - G % span) return b;
    return ((i / C) * X + x) * C + i % C;
}

struct FusedArgs {
    DevMesh m;
    NodeTab tab;
    double* vflux;         // [D][L]
    const int2* ekr;       // [D] wet level range of the edge [lo, hi)
    const double* uv;
    const double* helem;
    const double* w_e;
    const double* hnode;
    const double* hnn;
    const double* Z3d;     // optional
    const double* zbar3;   // optional
    double dt;
    int remap;             // block -> task remap, XCD count | chunk << 8 (1 = identity)
    int K;                 // tracers in this launch (<= FCT_KCHUNK)
    const double* T[FCT_KCHUNK];
    const double* Told[FCT_KCHUNK];
    double* Tnew[FCT_KCHUNK];
    double* LO;            // [K][N][L]
    double2* gnode;        // [K][N][L]
    double* adfh;          // [K][D][L]
    double* pf;            // [K][N][L]
    double* adfv;          // [K][N][L] adf_v at each node interface, from pass 2 for pass 3
};

// Wavefront -> task index. Single-instance launches number the wavefronts of the
// (remapped) grid; batched launches look up (instance, first wavefront) per workgroup.
__device__ __forceinline__ bool single_wave(int64_t total, int remap, int& w) {
    const int64_t ww = ((int64_t)block_tile(remap) * blockDim.x + threadIdx.x) >> 6;
    if (ww >= total) return false;
    w = __builtin_amdgcn_readfirstlane((int)ww);
    return true;
}

// Memory is accessed through global-address-space pointers (G). Pointers that a
// batched kernel reads from its descriptor array are generic to the compiler, which
// would otherwise emit flat loads for them.
template <class T>
using gptr = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gptr<T> G(T* p) { return (gptr<T>)p; }

// HIP's vector types are classes; load and store them as clang vectors of the same size.
typedef int fct_iv2 __attribute__((ext_vector_type(2)));
typedef int fct_iv4 __attribute__((ext_vector_type(4)));
typedef double fct_dv2 __attribute__((ext_vector_type(2)));
typedef double fct_dv4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ int2 ldv(const int2* p, size_t i) {
    return __builtin_bit_cast(int2, ((gptr<const fct_iv2>)p)[i]);
}
__device__ __forceinline__ int4 ldv(const int4* p, size_t i) {
    return __builtin_bit_cast(int4, ((gptr<const fct_iv4>)p)[i]);
}
__device__ __forceinline__ double2 ldv(const double2* p, size_t i) {
    return __builtin_bit_cast(double2, ((gptr<const fct_dv2>)p)[i]);
}
__device__ __forceinline__ double4 ldv(const double4* p, size_t i) {
    return __builtin_bit_cast(double4, ((gptr<const fct_dv4>)p)[i]);
}
// Loads at wave-uniform addresses of data that no kernel writes (mesh tables), through
// the constant address space so that the compiler issues scalar loads: they go through
// the scalar cache and leave the vector memory pipeline to the column gathers.
template <class T>
using cptr = __attribute__((address_space(4))) const T*;
__device__ __forceinline__ int ldc(const int* p, size_t i) { return ((cptr<int>)p)[i]; }
__device__ __forceinline__ double ldc(const double* p, size_t i) { return ((cptr<double>)p)[i]; }
__device__ __forceinline__ int2 ldc(const int2* p, size_t i) {
    return __builtin_bit_cast(int2, ((cptr<fct_iv2>)p)[i]);
}
__device__ __forceinline__ int4 ldc(const int4* p, size_t i) {
    return __builtin_bit_cast(int4, ((cptr<fct_iv4>)p)[i]);
}
__device__ __forceinline__ double2 ldc(const double2* p, size_t i) {
    return __builtin_bit_cast(double2, ((cptr<fct_dv2>)p)[i]);
}
__device__ __forceinline__ double4 ldc(const double4* p, size_t i) {
    return __builtin_bit_cast(double4, ((cptr<fct_dv4>)p)[i]);
}
__device__ __forceinline__ void stv(double2* p, size_t i, double2 v) {
    ((gptr<fct_dv2>)p)[i] = __builtin_bit_cast(fct_dv2, v);
}

__device__ __forceinline__ int batch_wave(const int2* __restrict__ blk, int remap, int& inst) {
    const int2 bi = blk[block_tile(remap)];
    inst = __builtin_amdgcn_readfirstlane(bi.x);
    return __builtin_amdgcn_readfirstlane(bi.y + (int)(threadIdx.x >> 6));
}

// ---------------------------------------------------------------------------
// vflux on [kmin_e, kmax_e) for EPW consecutive edges, loads issued together
// ---------------------------------------------------------------------------
template <int NS, int EPW>
__device__ __forceinline__ void vflux_group(const FusedArgs& a, int g) {
    const DevMesh& m = a.m;
    const int L = m.L;
    int e[EPW], tr[EPW];
    int2 et[EPW], kr[EPW];
    double4 c[EPW];
    int kmax = 0;
#pragma unroll
    for (int q = 0; q < EPW; ++q) {
        const int eq = g * EPW + q;
        e[q] = eq < m.D ? (a.tab.vorder ? ldc(a.tab.vorder, eq) : eq) : m.D - 1;
        et[q] = ldc(m.edge_tri, e[q]);
        kr[q] = eq < m.D ? ldc(a.ekr, e[q]) : make_int2(0, 0);
        c[q] = ldc(m.cross, e[q]);
        tr[q] = ldc(m.etr, e[q]);
        kmax = kr[q].y > kmax ? kr[q].y : kmax;
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        if (64 * s >= kmax) break;
        const int k = 64 * s + lane_id();
        double2 u1[EPW], u2[EPW];
        double h1[EPW], h2[EPW];
        if (k < kmax) {
#pragma unroll
            for (int q = 0; q < EPW; ++q) {
                const size_t i1 = (size_t)(et[q].x >= 0 ? et[q].x : 0) * L + k;
                const size_t i2 = (size_t)(et[q].y >= 0 ? et[q].y : 0) * L + k;
                u1[q] = ldv(reinterpret_cast<const double2*>(a.uv), i1);
                u2[q] = ldv(reinterpret_cast<const double2*>(a.uv), i2);
                h1[q] = G(a.helem)[i1];
                h2[q] = G(a.helem)[i2];
            }
        }
#pragma unroll
        for (int q = 0; q < EPW; ++q) {
            if (k < kr[q].x || k >= kr[q].y) continue;
            double t1 = 0.0, t2 = 0.0;
            if (in_bytes(tr[q], 0, k))
                t1 = (u1[q].x * c[q].y - u1[q].y * c[q].x) * h1[q];
            if (in_bytes(tr[q], 16, k))
                t2 = (u2[q].y * c[q].z - u2[q].x * c[q].w) * h2[q];
            G(a.vflux)[(size_t)e[q] * L + k] = t1 + t2;
        }
    }
}

// ---------------------------------------------------------------------------
// pass 1 of node n, tracer j: LO and gnode
// ---------------------------------------------------------------------------
template <int NS, int DM>
__device__ __forceinline__ void pass1_body(const FusedArgs& a, int n, int j, int4 hd) {
    const DevMesh& m = a.m;
    const int L = m.L;
    const int2 nr = make_int2(hd.x, hd.y), cnt = make_int2(hd.z & 0xff, hd.w);   // levels, counts
    // levels at which the adf_h pass reads this node's gnode, as a range [glo, ghi)
    const int ghz = hd.z;
    const int glo = (ghz >> 8) & 0xff, ghi = (ghz >> 16) & 0xff;
    const auto T = G(a.T[j]);
    const size_t b0 = (size_t)n * L;
    const size_t jo = (size_t)j * m.N * L;
    const int kneed = nr.y + 1 < L ? nr.y + 1 : L;
    int4 es[DM];
    int4 el[DM];
#pragma unroll
    for (int q = 0; q < DM; ++q) {
        es[q] = ldc(a.tab.edge, (size_t)n * DM + q);
        el[q] = ldc(a.tab.elem, (size_t)n * DM + q);
    }
    // per wavefront: the neighbour columns at the lane's level, then the node's own
    __shared__ double s_nb[4][DM + 1][64];
    double* nb = &s_nb[threadIdx.x >> 6][0][lane_id()];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        if (64 * s > nr.y) break;
        const int k = 64 * s + lane_id();
        const bool rd = k < kneed;
        const bool wet = in_layer(nr, k);
        const bool gslot = glo < ghi && glo < 64 * s + 64 && ghi > 64 * s;
        // own column, k-1 and k+1 of T for the vertical fluxes (the one-slot body takes
        // k-1 and k+1 from the neighbouring lanes below)
        const double Tk = rd ? T[b0 + k] : 0.0;
        double Tm = 0.0, Tp = 0.0;
        if (NS > 1) {
            Tm = (rd && k > 0) ? T[b0 + k - 1] : Tk;
            Tp = (k + 1 < kneed) ? T[b0 + k + 1] : 0.0;
        }
        // all neighbour columns and edge fluxes at once (one basic block: slot indices
        // past the node's count repeat slot 0 and are ignored when summing)
        double To[DM] = {}, vf[DM] = {};
        // the node's own fields at k (and k+1 for the lower interface), with the gathers
        double w0 = 0.0, a0 = 0.0, w1 = 0.0, a1 = 0.0, av = 0.0, hn = 0.0, hnd = 0.0;
        const size_t i = b0 + k, ip = b0 + (k + 1 < L ? k + 1 : k);
        if (wet) {
#pragma unroll
            for (int q = 0; q < DM; ++q) {
                vf[q] = G(a.vflux)[(size_t)(es[q].x >> 1) * L + k];
                To[q] = T[(size_t)es[q].y * L + k];
            }
            const double2 aa = ldv(m.aav, i);
            w0 = G(a.w_e)[i]; a0 = aa.x; av = aa.y;
            if (NS > 1) { w1 = G(a.w_e)[ip]; a1 = G(m.area)[ip]; }
            hn = G(a.hnn)[i]; hnd = G(a.hnode)[i];
            // the element vertices are the node and edge neighbours: the gradient below
            // reads them back from the lane's own LDS row by slot code (bits 16.. of the
            // element-slot word: 4 bits per vertex, DM = the node) instead of loading
            // them again
            if (gslot) {
#pragma unroll
                for (int q = 0; q < DM; ++q) nb[q * 64] = To[q];
                nb[DM * 64] = Tk;
            }
        }
        if (NS == 1) {
            // all lanes active here. Lanes past the loaded range hold 0, which is what
            // the two-slot body reads there; lane 63 is never used for k + 1 in one slot.
            // Tm is read on wet lanes only, where it is T[k-1], or T[0] at k = 0
            Tm = wave_prev(Tk, Tk);
            Tp = wave_next(Tk, 0.0);
            w1 = wave_next(w0, 0.0);
            a1 = wave_next(a0, 0.0);
        }
        if (!wet) continue;
        // LO: edges in reference order. Skipped terms add +0.0, which leaves a sum that
        // starts at +0.0 unchanged, so selects replace branches bit for bit.
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < DM; ++q) {
            const double v = vf[q];
            const bool side = es[q].x & 1;
            const double T1 = side ? To[q] : Tk;
            const double T2 = side ? Tk : To[q];
            const double av = fabs(v);
            const double f = -0.5 * (T1 * (v + av) + T2 * (v - av));
            const bool use = q < cnt.x && v != 0.0;
            acc = acc + (use ? (side ? -f : f) : 0.0);
        }
        double fv0, fv1 = 0.0;
        {
            const double aw = fabs(w0);
            fv0 = (-0.5 * (Tk * (w0 + aw) + Tm * (w0 - aw))) * a0;
        }
        if (in_layer(nr, k + 1)) {
            const double aw = fabs(w1);
            fv1 = (-0.5 * (Tp * (w1 + aw) + Tk * (w1 - aw))) * a1;
        }
        const double safe_a = av > 0.0 ? av : 1.0, safe_hn = hn > 0.0 ? hn : 1.0;
        const double numer = Tk * hnd + (acc + (fv0 - fv1)) * a.dt / safe_a;
        const double lo = numer / safe_hn;
        G(a.LO)[jo + i] = (hn > 0.0 && av > 0.0) ? lo : 0.0;
        // gnode: elements in increasing id, only on the levels the adf_h pass reads
        if (!gslot) continue;
        double nx = 0.0, ny = 0.0, den = 0.0;
#pragma unroll
        for (int q = 0; q < DM; ++q) {
            const int e = el[q].x;
            const bool use = q < cnt.y && in_bytes(es[q].w, 0, k);
            const size_t g6 = 6 * (size_t)e;
            const unsigned codes = (unsigned)es[q].w >> 16;
            const double tv0 = nb[(codes & 0xf) * 64], tv1 = nb[((codes >> 4) & 0xf) * 64];
            const double tv2 = nb[((codes >> 8) & 0xf) * 64];
            const double gx = ldc(m.gs, g6) * tv0 + ldc(m.gs, g6 + 1) * tv1 + ldc(m.gs, g6 + 2) * tv2;
            const double gy = ldc(m.gs, g6 + 3) * tv0 + ldc(m.gs, g6 + 4) * tv1 + ldc(m.gs, g6 + 5) * tv2;
            const double ea = ldc(m.earea, e);
            nx = nx + (use ? gx * ea : 0.0);
            ny = ny + (use ? gy * ea : 0.0);
            den = den + (use ? ea : 0.0);
        }
        const double d = den > 0.0 ? den : 1.0;
        if (k >= glo && k < ghi) stv(a.gnode, jo + i, make_double2(nx / d, ny / d));
    }
}

// Pass 1 of a shallow node (wet layers end above level 64, one 64-level slot) for its KT
// tracers in one wavefront: the tracer-independent loads (the edge fluxes, w, area,
// areasvol, hnode, hnode_new) are made once; per tracer the arithmetic is pass1_body's.
// KT is a compile-time count so that the tracer pointers are read at constant offsets.
template <int DM, int KT>
__device__ __forceinline__ void pass1_shallow_all(const FusedArgs& a, int n, int4 hd, int j0) {
    const DevMesh& m = a.m;
    const int L = m.L;
    const int2 nr = make_int2(hd.x, hd.y), cnt = make_int2(hd.z & 0xff, hd.w);
    const int glo = (hd.z >> 8) & 0xff, ghi = (hd.z >> 16) & 0xff;
    const size_t b0 = (size_t)n * L;
    const int kneed = nr.y + 1 < L ? nr.y + 1 : L;
    int4 es[DM];
    int4 el[DM];
#pragma unroll
    for (int q = 0; q < DM; ++q) {
        es[q] = ldc(a.tab.edge, (size_t)n * DM + q);
        el[q] = ldc(a.tab.elem, (size_t)n * DM + q);
    }
    __shared__ double s_nb[4][DM + 1][64];
    double* nb = &s_nb[threadIdx.x >> 6][0][lane_id()];
    const int k = lane_id();
    const bool rd = k < kneed;
    const bool wet = in_layer(nr, k);
    const bool gslot = glo < ghi;
    const size_t i = b0 + k;
    double vf[DM] = {};
    double w0 = 0.0, a0 = 0.0, av = 0.0, hn = 0.0, hnd = 0.0;
    if (wet) {
#pragma unroll
        for (int q = 0; q < DM; ++q) vf[q] = G(a.vflux)[(size_t)(es[q].x >> 1) * L + k];
        const double2 aa = ldv(m.aav, i);
        w0 = G(a.w_e)[i]; a0 = aa.x; av = aa.y;
        hn = G(a.hnn)[i]; hnd = G(a.hnode)[i];
    }
    const double w1 = wave_next(w0, 0.0), a1 = wave_next(a0, 0.0);
#pragma unroll
    for (int jj = 0; jj < KT; ++jj) {
        const int j = j0 + jj;   // tracer slot
        const auto T = G(a.T[j]);
        const size_t jo = (size_t)j * m.N * L;
        const double Tk = rd ? T[b0 + k] : 0.0;
        double To[DM] = {};
        if (wet) {
#pragma unroll
            for (int q = 0; q < DM; ++q) To[q] = T[(size_t)es[q].y * L + k];
            if (gslot) {
#pragma unroll
                for (int q = 0; q < DM; ++q) nb[q * 64] = To[q];
                nb[DM * 64] = Tk;
            }
        }
        const double Tm = wave_prev(Tk, Tk);
        const double Tp = wave_next(Tk, 0.0);
        if (!wet) continue;
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < DM; ++q) {
            const double v = vf[q];
            const bool side = es[q].x & 1;
            const double T1 = side ? To[q] : Tk;
            const double T2 = side ? Tk : To[q];
            const double aq = fabs(v);
            const double f = -0.5 * (T1 * (v + aq) + T2 * (v - aq));
            const bool use = q < cnt.x && v != 0.0;
            acc = acc + (use ? (side ? -f : f) : 0.0);
        }
        double fv0, fv1 = 0.0;
        {
            const double aw = fabs(w0);
            fv0 = (-0.5 * (Tk * (w0 + aw) + Tm * (w0 - aw))) * a0;
        }
        if (in_layer(nr, k + 1)) {
            const double aw = fabs(w1);
            fv1 = (-0.5 * (Tp * (w1 + aw) + Tk * (w1 - aw))) * a1;
        }
        const double safe_a = av > 0.0 ? av : 1.0, safe_hn = hn > 0.0 ? hn : 1.0;
        const double numer = Tk * hnd + (acc + (fv0 - fv1)) * a.dt / safe_a;
        const double lo = numer / safe_hn;
        G(a.LO)[jo + i] = (hn > 0.0 && av > 0.0) ? lo : 0.0;
        if (!gslot) continue;
        double nx = 0.0, ny = 0.0, den = 0.0;
#pragma unroll
        for (int q = 0; q < DM; ++q) {
            const int e = el[q].x;
            const bool use = q < cnt.y && in_bytes(es[q].w, 0, k);
            const size_t g6 = 6 * (size_t)e;
            const unsigned codes = (unsigned)es[q].w >> 16;
            const double tv0 = nb[(codes & 0xf) * 64], tv1 = nb[((codes >> 4) & 0xf) * 64];
            const double tv2 = nb[((codes >> 8) & 0xf) * 64];
            const double gx = ldc(m.gs, g6) * tv0 + ldc(m.gs, g6 + 1) * tv1 + ldc(m.gs, g6 + 2) * tv2;
            const double gy = ldc(m.gs, g6 + 3) * tv0 + ldc(m.gs, g6 + 4) * tv1 + ldc(m.gs, g6 + 5) * tv2;
            const double ea = ldc(m.earea, e);
            nx = nx + (use ? gx * ea : 0.0);
            ny = ny + (use ? gy * ea : 0.0);
            den = den + (use ? ea : 0.0);
        }
        const double d = den > 0.0 ? den : 1.0;
        if (k >= glo && k < ghi) stv(a.gnode, jo + i, make_double2(nx / d, ny / d));
    }
}

template <int NS, int DM>
__device__ __forceinline__ void pass1_task(const FusedArgs& a, int n, int j) {
    pass1_body<NS, DM>(a, n, j, ldc(a.tab.hdr, n));
}

// ---------------------------------------------------------------------------
// adf_h(e, k) for EPW consecutive edges, tracer j
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 grad_of(const double* gs, int el, double t0, double t1,
                                           double t2) {
    const size_t g6 = 6 * (size_t)el;
    double2 g;
    g.x = ldc(gs, g6) * t0 + ldc(gs, g6 + 1) * t1 + ldc(gs, g6 + 2) * t2;
    g.y = ldc(gs, g6 + 3) * t0 + ldc(gs, g6 + 4) * t1 + ldc(gs, g6 + 5) * t2;
    return g;
}

template <int NS, int EPW>
__device__ __forceinline__ void adfh_group(const FusedArgs& a, int g, int j) {
    const DevMesh& m = a.m;
    const int L = m.L;
    const auto T = G(a.T[j]);
    const auto To = G(a.Told[j]);
    const double2* gn = a.gnode + (size_t)j * m.N * L;
    const auto out = G(a.adfh + (size_t)j * m.D * L);
    int e[EPW], u[EPW][2], d[EPW][2], cu[EPW], cd[EPW];
    int2 kr[EPW], nd[EPW], ur[EPW], ud[EPW];
    double2 cc[EPW];
    int kmax = 0;
#pragma unroll
    for (int q = 0; q < EPW; ++q) {
        const int eq = g * EPW + q;
        e[q] = eq < m.D ? eq : m.D - 1;
        kr[q] = eq < m.D ? ldc(a.ekr, e[q]) : make_int2(0, 0);
        nd[q] = ldc(m.edges, e[q]);
        ur[q] = ldc(m.utri_rng, e[q]);
        ud[q] = ldc(m.updn, e[q]);
        cc[q] = ldc(m.edge_c, e[q]);
        const int4 va = ldc(m.udv, 2 * (size_t)e[q]);
        const int vb = ldc(reinterpret_cast<const int*>(m.udv), 8 * (size_t)e[q] + 4);
        u[q][0] = va.x; u[q][1] = va.y; d[q][0] = va.z; d[q][1] = va.w;
        cu[q] = vb & 3; cd[q] = (vb >> 2) & 3;
    }
#pragma unroll
    for (int q = 0; q < EPW; ++q) kmax = kr[q].y > kmax ? kr[q].y : kmax;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        if (64 * s >= kmax) break;
        const int k = 64 * s + lane_id();
        double vf[EPW] = {}, T1[EPW] = {}, T2[EPW] = {}, O1[EPW] = {}, O2[EPW] = {};
        double tu[EPW][3] = {}, td[EPW][3] = {};
        // per edge: does any wet level of this slot use the up/dn triangle gradients
        // (nt), does any use the node gradients gnode (ng; few edges, so those loads
        // wait until the edge is computed and take no registers before)
        bool nt[EPW], ng[EPW];
#pragma unroll
        for (int q = 0; q < EPW; ++q) {
            const int lo = kr[q].x > 64 * s ? kr[q].x : 64 * s;
            const int hi = kr[q].y < 64 * s + 64 ? kr[q].y : 64 * s + 64;
            const int tl = lo > ur[q].x ? lo : ur[q].x, th = hi < ur[q].y ? hi : ur[q].y;
            nt[q] = tl < th;
            ng[q] = lo < hi && !(tl == lo && th == hi);
        }
        if (k < kmax) {
#pragma unroll
            for (int q = 0; q < EPW; ++q) {
                const size_t i1 = (size_t)nd[q].x * L + k, i2 = (size_t)nd[q].y * L + k;
                vf[q] = G(a.vflux)[(size_t)e[q] * L + k];
                T1[q] = T[i1]; T2[q] = T[i2]; O1[q] = To[i1]; O2[q] = To[i2];
                if (nt[q]) {
                    // the up triangle's vertex at slot cu is the edge's first node, the down
                    // triangle's at slot cd its second: their T is already loaded
                    const double a0 = T[(size_t)u[q][0] * L + k], a1 = T[(size_t)u[q][1] * L + k];
                    const double b0 = T[(size_t)d[q][0] * L + k], b1 = T[(size_t)d[q][1] * L + k];
                    tu[q][0] = cu[q] == 0 ? T1[q] : a0;
                    tu[q][1] = cu[q] == 0 ? a0 : (cu[q] == 1 ? T1[q] : a1);
                    tu[q][2] = cu[q] == 2 ? T1[q] : a1;
                    td[q][0] = cd[q] == 0 ? T2[q] : b0;
                    td[q][1] = cd[q] == 0 ? b0 : (cd[q] == 1 ? T2[q] : b1);
                    td[q][2] = cd[q] == 2 ? T2[q] : b1;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < EPW; ++q) {
            if (k < kr[q].x || k >= kr[q].y) continue;
            const bool tri = k >= ur[q].x && k < ur[q].y;
            const double A1 = ab2(O1[q], T1[q]), A2 = ab2(O2[q], T2[q]);
            double2 gu = make_double2(0.0, 0.0), gd = gu;
            if (nt[q]) {
                gu = grad_of(m.gs, ud[q].x, tu[q][0], tu[q][1], tu[q][2]);
                gd = grad_of(m.gs, ud[q].y, td[q][0], td[q][1], td[q][2]);
            }
            double2 ga = make_double2(0.0, 0.0), gb = ga;
            if (ng[q]) {
                ga = ldv(gn, (size_t)nd[q].x * L + k);
                gb = ldv(gn, (size_t)nd[q].y * L + k);
            }
            const double g0 = tri ? gu.x : ga.x, g1 = tri ? gd.x : gb.x;
            const double g2 = tri ? gu.y : ga.y, g3 = tri ? gd.y : gb.y;
            const double v = vf[q];
            const double av = fabs(v);
            const double flo = -0.5 * (T1[q] * (v + av) + T2[q] * (v - av));
            const double diff = A2 - A1;
            const double Tm1 = A1 + (2.0 * diff + cc[q].x * g0 + cc[q].y * g2) / 6.0;
            const double Tm2 = A2 - (2.0 * diff + cc[q].x * g1 + cc[q].y * g3) / 6.0;
            const double ho = -0.5 * ((v + av) * Tm1 + (v - av) * Tm2);
            const double r = v != 0.0 ? ho - flo : 0.0;
            out[(size_t)e[q] * L + k] = r;
        }
    }
}

// ---------------------------------------------------------------------------
// adf_h over edge patches: KT tracers, KC levels per chunk, TPE threads per edge
// ---------------------------------------------------------------------------
// A workgroup takes one edge patch (the edges a compact node patch owns). For each chunk
// of KC levels it loads (T, ttfAB) of the patch's node list into LDS, KC lanes per node
// column so that every load instruction reads whole 8*KC-byte row segments, and the
// patch's vflux likewise; then each edge's TPE threads compute its chunk levels from
// LDS (the same expressions in the same order as adfh_group, so the same bits) and leave
// adf_h in LDS, from where it is stored as row segments on the edge's wet levels. The
// loads of the next chunk are issued before the current chunk is computed, so their
// latency overlaps the arithmetic. Each node column of the patch is read from memory
// once per chunk instead of once per incident edge; only the node-averaged gradients
// (read on the few levels where an edge does not use its up/down triangles) are still
// gathered per edge.
template <int KT, int KC, int TPE>
__device__ __forceinline__ void adfh_patch(const FusedArgs& a, int p, int j0) {
    constexpr int NT = FCT_EP * TPE;              // threads
    constexpr int PS = FCT_PU + 1;                // row pitches padded against bank conflicts
    constexpr int ES = FCT_EP + 1;
    constexpr int NI = (FCT_PU * KC + NT - 1) / NT;   // node-column items per thread
    constexpr int NV = (FCT_EP * KC + NT - 1) / NT;   // edge-column items per thread
    constexpr int LPT = KC / TPE;                 // chunk levels per thread
    __shared__ double2 ta[KT][KC][PS];            // (T, ttfAB) of the node list
    __shared__ double vo[KT][KC][ES];             // vflux in (tracer 0's rows), adf_h out
    __shared__ int2 ekr_s[FCT_EP];
    const DevMesh& m = a.m;
    const int L = m.L;
    const int4 pt = ldc(a.tab.ep.ptab, p);
    const int first = pt.x, noff = pt.y, cnt = pt.z & 0xffff, U = pt.z >> 16;
    const int klo = pt.w & 0xff, khi = (pt.w >> 8) & 0xff;
    const int t = threadIdx.x, q = t % FCT_EP, kq = (t / FCT_EP) * LPT;
    int2 kr = make_int2(0, 0), ur = kr, nd = kr, sl = kr;
    double2 cc = make_double2(0.0, 0.0);
    double gu[6] = {}, gd[6] = {};
    if (q < cnt) {
        const int e = G(a.tab.ep.elist)[first + q];
        kr = ldv(a.ekr, e);
        ur = ldv(m.utri_rng, e);
        nd = ldv(m.edges, e);
        sl = ldv(a.tab.ep.eslot, (size_t)first + q);
        cc = ldv(m.edge_c, e);
        if (ur.x < ur.y) {
            const int2 ud = ldv(m.updn, e);
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                gu[i] = G(m.gs)[6 * (size_t)ud.x + i];
                gd[i] = G(m.gs)[6 * (size_t)ud.y + i];
            }
        }
        if (t < FCT_EP) ekr_s[q] = kr;
    }
    const int s1 = sl.x & 0xff, s2 = (sl.x >> 8) & 0xff;
    const int sua = (sl.x >> 16) & 0xff, sub = (sl.x >> 24) & 0xff;
    const int sda = sl.y & 0xff, sdb = (sl.y >> 8) & 0xff;
    const int cu = (sl.y >> 16) & 3, cd = (sl.y >> 18) & 3;
    // row offsets of this thread's items: node columns (slot s, level kk) and edge
    // columns (edge q, level kk), i = t + r * NT, s or q = i / KC, kk = i % KC
    size_t nrow[NI], erow[NV];
#pragma unroll
    for (int r = 0; r < NI; ++r) {
        const int i = t + r * NT, s = i / KC;
        nrow[r] = s < U ? (size_t)G(a.tab.ep.pnodes)[noff + s] * L : 0;
    }
#pragma unroll
    for (int r = 0; r < NV; ++r) {
        const int i = t + r * NT, qq = i / KC;
        erow[r] = qq < cnt ? (size_t)G(a.tab.ep.elist)[first + qq] * L : 0;
    }
    double pT[NI][KT], pO[NI][KT], pv[NV];
#define FCT_PREFETCH(K0)                                                                    \
    do {                                                                                    \
        _Pragma("unroll") for (int r = 0; r < NI; ++r) {                                    \
            const int i = t + r * NT, s = i / KC, k = (K0) + i % KC;                        \
            const bool ok = s < U && k < L;                                                 \
            _Pragma("unroll") for (int jj = 0; jj < KT; ++jj) {                             \
                pT[r][jj] = ok ? G(a.T[j0 + jj])[nrow[r] + k] : 0.0;                        \
                pO[r][jj] = ok ? G(a.Told[j0 + jj])[nrow[r] + k] : 0.0;                     \
            }                                                                               \
        }                                                                                   \
        _Pragma("unroll") for (int r = 0; r < NV; ++r) {                                    \
            const int i = t + r * NT, qq = i / KC, k = (K0) + i % KC;                       \
            pv[r] = (qq < cnt && k < L) ? G(a.vflux)[erow[r] + k] : 0.0;                    \
        }                                                                                   \
    } while (0)
    FCT_PREFETCH(klo);
    for (int k0 = klo; k0 < khi; k0 += KC) {
#pragma unroll
        for (int r = 0; r < NI; ++r) {
            const int i = t + r * NT, s = i / KC, kk = i % KC;
            if (s < U) {
#pragma unroll
                for (int jj = 0; jj < KT; ++jj)
                    ta[jj][kk][s] = make_double2(pT[r][jj], ab2(pO[r][jj], pT[r][jj]));
            }
        }
#pragma unroll
        for (int r = 0; r < NV; ++r) {
            const int i = t + r * NT, qq = i / KC, kk = i % KC;
            if (qq < cnt) vo[0][kk][qq] = pv[r];
        }
        __syncthreads();
        if (k0 + KC < khi) FCT_PREFETCH(k0 + KC);
        if (q < cnt) {
#pragma unroll
            for (int kl = 0; kl < LPT; ++kl) {
                const int kk = kq + kl, k = k0 + kk;
                const double v = vo[0][kk][q];
                const bool wet = k >= kr.x && k < kr.y;
                const bool tri = k >= ur.x && k < ur.y;
#pragma unroll
                for (int jj = 0; jj < KT; ++jj) {
                    double r = 0.0;
                    if (wet) {
                        const double2 e1 = ta[jj][kk][s1], e2 = ta[jj][kk][s2];
                        const double T1 = e1.x, T2 = e2.x, A1 = e1.y, A2 = e2.y;
                        double g0, g1, g2, g3;
                        if (tri) {
                            const double a0 = ta[jj][kk][sua].x, a1 = ta[jj][kk][sub].x;
                            const double b0 = ta[jj][kk][sda].x, b1 = ta[jj][kk][sdb].x;
                            const double u0 = cu == 0 ? T1 : a0;
                            const double u1 = cu == 0 ? a0 : (cu == 1 ? T1 : a1);
                            const double u2 = cu == 2 ? T1 : a1;
                            const double d0 = cd == 0 ? T2 : b0;
                            const double d1 = cd == 0 ? b0 : (cd == 1 ? T2 : b1);
                            const double d2 = cd == 2 ? T2 : b1;
                            g0 = gu[0] * u0 + gu[1] * u1 + gu[2] * u2;
                            g2 = gu[3] * u0 + gu[4] * u1 + gu[5] * u2;
                            g1 = gd[0] * d0 + gd[1] * d1 + gd[2] * d2;
                            g3 = gd[3] * d0 + gd[4] * d1 + gd[5] * d2;
                        } else {
                            const double2* gn = a.gnode + (size_t)(j0 + jj) * m.N * L;
                            const double2 ga = ldv(gn, (size_t)nd.x * L + k);
                            const double2 gb = ldv(gn, (size_t)nd.y * L + k);
                            g0 = ga.x; g1 = gb.x; g2 = ga.y; g3 = gb.y;
                        }
                        const double av = fabs(v);
                        const double flo = -0.5 * (T1 * (v + av) + T2 * (v - av));
                        const double diff = A2 - A1;
                        const double Tm1 = A1 + (2.0 * diff + cc.x * g0 + cc.y * g2) / 6.0;
                        const double Tm2 = A2 - (2.0 * diff + cc.x * g1 + cc.y * g3) / 6.0;
                        const double ho = -0.5 * ((v + av) * Tm1 + (v - av) * Tm2);
                        r = v != 0.0 ? ho - flo : 0.0;
                    }
                    vo[jj][kk][q] = r;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < NV; ++r) {
            const int i = t + r * NT, qq = i / KC, kk = i % KC, k = k0 + kk;
            if (qq < cnt) {
                const int2 rr = ekr_s[qq];
                if (k >= rr.x && k < rr.y) {
#pragma unroll
                    for (int jj = 0; jj < KT; ++jj)
                        G(a.adfh + (size_t)(j0 + jj) * m.D * L)[erow[r] + k] = vo[jj][kk][qq];
                }
            }
        }
        __syncthreads();
    }
#undef FCT_PREFETCH
}

// ---------------------------------------------------------------------------
// vertical antidiffusive flux (adv_flux_ver_ho - adv_flux_ver, tracer_adv.py:241-295,
// 108-120) at interface k from the column's ttfAB at k-2..k+1 and T at k-1, k
// ---------------------------------------------------------------------------
// The three QR4C gradients of interface k are one quotient at three levels:
// qc(k) = (ttf[k-1] - ttf[k]) / nz_safe(Z[k-1] - Z[k]), qu(k) = qc(k+1), qd(k) = qc(k-1)
// (the same operands in the same order, so the same bits): the column divides once per
// level and takes qu, qd by lane shifts.
__device__ __forceinline__ double adfv_at(int k, int2 nr, double A1, double A, double Tm,
                                          double Tk, double W, double ar, double Znz,
                                          double Zm1, double qc, double qu, double qd,
                                          double zb) {
    double hi = 0.0;
    if (in_iface(nr, k)) {
        if (k == nr.x) {
            hi = -A * W * ar;
        } else if (k == nr.y) {
            hi = 0.0;
        } else if (k == nr.x + 1 || k == nr.y - 1) {
            hi = -0.5 * (A1 + A) * W * ar;
        } else if (k >= nr.x + 2 && k <= nr.y - 2) {
            const double Tmean1 = A + (2.0 * qc + qu) * (zb - Znz) / 3.0;
            const double Tmean2 = A1 + (2.0 * qc + qd) * (zb - Zm1) / 3.0;
            hi = -0.5 * (Tmean1 + Tmean2) * W * ar;
        }
    }
    double lo = 0.0;
    if (in_layer(nr, k)) {
        const double aw = fabs(W);
        lo = (-0.5 * (Tk * (W + aw) + Tm * (W - aw))) * ar;
    }
    return hi - lo;
}

// Column of adf_v for the node: every lane k <= nzmax (0 elsewhere). All lanes active.
// Also hands back areasvol at every level (loaded with the area, clamped like it).
// The column's w, area, areasvol, depth Z at k and interface depth at every lane
// (clamped below the last level): the tracer-independent inputs of adfv_column.
template <int NS>
__device__ __forceinline__ void adfv_inputs(const FusedArgs& a, int n, double (&W)[NS],
                                            double (&ar)[NS], double (&avc)[NS],
                                            double (&Zc)[NS], double (&zb)[NS]) {
    const DevMesh& m = a.m;
    const int L = m.L;
    const size_t b0 = (size_t)n * L;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int k = 64 * s + lane_id();
        const int kc = k < L ? k : L - 1;
        const size_t i = b0 + kc;
        W[s] = G(a.w_e)[i];
        const double2 aa = ldv(m.aav, i);
        ar[s] = aa.x;
        avc[s] = aa.y;
        if (a.Z3d) {
            Zc[s] = G(a.Z3d)[i];
            zb[s] = G(a.zbar3)[i];
        } else {
            const double2 z = ldv(m.zzb, kc);
            Zc[s] = z.x;
            zb[s] = z.y;
        }
    }
}

template <int NS>
__device__ __forceinline__ void adfv_compute(const FusedArgs& a, int2 nr,
                                             const double (&Tc)[NS], const double (&Ac)[NS],
                                             const double (&W)[NS], const double (&ar)[NS],
                                             const double (&Zc)[NS], const double (&zb)[NS],
                                             double (&out)[NS]) {
    const int L = a.m.L;
    double Tm[NS], A1[NS];
    col_prev<NS>(Tc, Tm, col_bcast0<NS>(Tc));
    col_prev<NS>(Ac, A1, col_bcast0<NS>(Ac));
    // The reference's stencil depths (_z_stencil, tracer_adv.py:228-238) clamp Z[k-1] and
    // Z[k-2] at the top and Z[k+1] at the bottom. qc is needed at levels nr.x+1 .. nr.y-1
    // only (qd at nr.x+2 reads nr.x+1, qu at nr.y-2 reads nr.y-1): there 1 <= k <= L-2,
    // no clamp applies, and Z[k-1] is the plain lane shift of Z.
    const double z0 = col_bcast0<NS>(Zc);
    double Zm1[NS], qc[NS], qu[NS], qd[NS];
    col_prev<NS>(Zc, Zm1, z0);
#pragma unroll
    for (int s = 0; s < NS; ++s)
        qc[s] = 64 * s <= nr.y ? (A1[s] - Ac[s]) / nz_safe(Zm1[s] - Zc[s]) : 0.0;
    col_next<NS>(qc, qu, 0.0);
    col_prev<NS>(qc, qd, 0.0);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int k = 64 * s + lane_id();
        out[s] = 0.0;
        if (k <= nr.y && k < L)
            out[s] = adfv_at(k, nr, A1[s], Ac[s], Tm[s], Tc[s], W[s], ar[s], Zc[s], Zm1[s],
                             qc[s], qu[s], qd[s], zb[s]);
    }
}

template <int NS>
__device__ __forceinline__ void adfv_column(const FusedArgs& a, int n, int2 nr,
                                            const double (&Tc)[NS], const double (&Ac)[NS],
                                            double (&out)[NS], double (&avc)[NS]) {
    double W[NS], ar[NS], Zc[NS], zb[NS];
    adfv_inputs<NS>(a, n, W, ar, avc, Zc, zb);
    adfv_compute<NS>(a, nr, Tc, Ac, W, ar, Zc, zb, out);
}

// ---------------------------------------------------------------------------
// pass 2 of node n, tracer j: cluster bounds, P+/P-, limiter factors
// ---------------------------------------------------------------------------
template <int NS, int DM>
__device__ __forceinline__ void pass2_body(const FusedArgs& a, int n, int j, int2 nr, int2 cnt) {
    const DevMesh& m = a.m;
    const int L = m.L;
    const size_t jo = (size_t)j * m.N * L;
    const size_t b0 = (size_t)n * L;
    const auto Tb = G(a.T[j]);
    const auto LOb = G(a.LO + jo);
    const auto Tr = G(Tb + b0);
    const auto Or = G(a.Told[j] + b0);
    const auto LOr = G(LOb + b0);
    const auto adfh = G(a.adfh + (size_t)j * m.D * L);
    const int kneed = nr.y + 1 < L ? nr.y + 1 : L;
    int4 es[DM];
#pragma unroll
    for (int q = 0; q < DM; ++q) es[q] = ldc(a.tab.edge, (size_t)n * DM + q);
    double Tc[NS], Ac[NS], LOc[NS], tvx[NS], tvn[NS], ap[NS], am[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        Tc[s] = 0.0; Ac[s] = 0.0; LOc[s] = 0.0;
        tvx[s] = -INFINITY; tvn[s] = INFINITY; ap[s] = 0.0; am[s] = 0.0;
        if (64 * s >= kneed) continue;
        const int k = 64 * s + lane_id();
        const bool rd = k < kneed;
        const bool wet = in_layer(nr, k);
        // own column and the neighbours' LO, T and the edges' antidiffusive fluxes,
        // loaded in one basic block
        double lnb[DM] = {}, tnb[DM] = {}, fh[DM] = {};
        double lo_own = 0.0;
        if (rd) {
            Tc[s] = Tr[k];
            Ac[s] = ab2(Or[k], Tc[s]);
            lo_own = LOr[k];
#pragma unroll
            for (int q = 0; q < DM; ++q) {
                const size_t r = (size_t)es[q].y * L + k;
                lnb[q] = LOb[r];
                tnb[q] = Tb[r];
                fh[q] = adfh[(size_t)(es[q].x >> 1) * L + k];
            }
        }
        LOc[s] = wet ? lo_own : 0.0;
        // a1-a3: max/min of (LO, T) over the vertices of the node's triangles that are
        // wet at k, and the -+1e3 pad when one of them is dry at k
        bool act = false, inact = false;
#pragma unroll
        for (int q = 0; q < DM; ++q) {
            const bool in = in_bytes(es[q].w, 0, k);
            act = act || (q < cnt.y && in);
            inact = inact || (q < cnt.y && !in);
        }
        double mx = act ? fmax(LOc[s], Tc[s]) : -INFINITY;
        double mn = act ? fmin(LOc[s], Tc[s]) : INFINITY;
#pragma unroll
        for (int q = 0; q < DM; ++q) {
            const bool inc = q < cnt.x && (in_bytes(es[q].z, 0, k) || in_bytes(es[q].z, 16, k));
            mx = inc ? fmax(mx, fmax(lnb[q], tnb[q])) : mx;
            mn = inc ? fmin(mn, fmin(lnb[q], tnb[q])) : mn;
        }
        mx = inact ? fmax(mx, -FCT_BIG) : mx;
        mn = inact ? fmin(mn, FCT_BIG) : mn;
        tvx[s] = rd ? mx : -INFINITY;
        tvn[s] = rd ? mn : INFINITY;
        // b1: P+/P- from the incident edges in reference order
#pragma unroll
        for (int q = 0; q < DM; ++q) {
            const double f = (es[q].x & 1) ? -fh[q] : fh[q];
            const bool use = wet && q < cnt.x;
            ap[s] = ap[s] + (use ? fmax(f, 0.0) : 0.0);
            am[s] = am[s] + (use ? fmin(f, 0.0) : 0.0);
        }
    }
    double tvxm[NS], tvxp[NS], tvnm[NS], tvnp[NS];
    col_prev<NS>(tvx, tvxm, col_bcast0<NS>(tvx));
    col_next<NS>(tvx, tvxp, 0.0);
    col_prev<NS>(tvn, tvnm, col_bcast0<NS>(tvn));
    col_next<NS>(tvn, tvnp, 0.0);
    double fvc[NS], fvb[NS], avc[NS];
    adfv_column<NS>(a, n, nr, Tc, Ac, fvc, avc);
    col_next<NS>(fvc, fvb, 0.0);
    // for pass 3, which reads it instead of computing it again
    const auto Fv = G(a.adfv + jo);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int k = 64 * s + lane_id();
        if (k <= nr.y && k < L) Fv[b0 + k] = fvc[s];
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int k = 64 * s + lane_id();
        if (!in_layer(nr, k)) continue;
        const size_t i = b0 + k;
        const bool use_own = (k == nr.x) || (k == nr.y - 1);
        const double cmax = use_own ? tvx[s] : fmax(fmax(tvxm[s], tvx[s]), tvxp[s]);
        const double cmin = use_own ? tvn[s] : fmin(fmin(tvnm[s], tvn[s]), tvnp[s]);
        const double ttf_max = cmax - LOc[s];
        const double ttf_min = cmin - LOc[s];
        const double pos_v = fmax(fvc[s], 0.0) + fmax(-fvb[s], 0.0);
        const double neg_v = fmin(fvc[s], 0.0) + fmin(-fvb[s], 0.0);
        const double fplus = pos_v + ap[s];
        const double fminus = neg_v + am[s];
        const double av = avc[s], hn = G(a.hnn)[i];
        const double safe_a = av > 0.0 ? av : 1.0, safe_hn = hn > 0.0 ? hn : 1.0;
        const double flux_pos = fplus * a.dt / safe_a / safe_hn + FCT_FLUX_EPS;
        const double flux_neg = fminus * a.dt / safe_a / safe_hn - FCT_FLUX_EPS;
        const bool valid = av > 0.0 && hn > 0.0;
        // fmin(1, x / y) is 1 wherever x / y >= 1 (rounding is monotone and 1 is a
        // double), which the signs and one comparison decide: only lanes where the
        // limiter may bind divide, and a wavefront with none skips the divisions
        const bool bp = valid && !(flux_pos > 0.0 && ttf_max >= flux_pos);
        const bool bm = valid && !(flux_neg < 0.0 && ttf_min <= flux_neg);
        double fp = 1.0, fm = 1.0;
        if (bp) fp = fmin(1.0, ttf_max / flux_pos);
        if (bm) fm = fmin(1.0, ttf_min / flux_neg);
        // (plus, minus) factor pairs, interleaved: pass 3 reads both with one load
        stv(reinterpret_cast<double2*>(a.pf), jo + i, make_double2(fp, fm));
    }
}

// Pass 2 of a shallow node (one 64-level slot) for its KT tracers in one wavefront: w,
// area, areasvol, the stencil depths and hnode_new are loaded once; per tracer the
// arithmetic is pass2_body's.
template <int DM, int KT>
__device__ __forceinline__ void pass2_shallow_all(const FusedArgs& a, int n, int2 nr, int2 cnt, int j0) {
    constexpr int NS = 1;
    const DevMesh& m = a.m;
    const int L = m.L;
    double Ws[NS], ars[NS], avs[NS], Zcs[NS], zbs[NS];
    adfv_inputs<NS>(a, n, Ws, ars, avs, Zcs, zbs);
    double hns[NS];
    {
        const int kl = lane_id();
        hns[0] = in_layer(nr, kl) ? G(a.hnn)[(size_t)n * L + kl] : 0.0;
    }
#pragma unroll
    for (int jj = 0; jj < KT; ++jj) {
        const int j = j0 + jj;   // tracer slot
        const size_t jo = (size_t)j * m.N * L;
        const size_t b0 = (size_t)n * L;
        const auto Tb = G(a.T[j]);
        const auto LOb = G(a.LO + jo);
        const auto Tr = G(Tb + b0);
        const auto Or = G(a.Told[j] + b0);
        const auto LOr = G(LOb + b0);
        const auto adfh = G(a.adfh + (size_t)j * m.D * L);
        const int kneed = nr.y + 1 < L ? nr.y + 1 : L;
        int4 es[DM];
#pragma unroll
        for (int q = 0; q < DM; ++q) es[q] = ldc(a.tab.edge, (size_t)n * DM + q);
        double Tc[NS], Ac[NS], LOc[NS], tvx[NS], tvn[NS], ap[NS], am[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            Tc[s] = 0.0; Ac[s] = 0.0; LOc[s] = 0.0;
            tvx[s] = -INFINITY; tvn[s] = INFINITY; ap[s] = 0.0; am[s] = 0.0;
            if (64 * s >= kneed) continue;
            const int k = 64 * s + lane_id();
            const bool rd = k < kneed;
            const bool wet = in_layer(nr, k);
            // own column and the neighbours' LO, T and the edges' antidiffusive fluxes,
            // loaded in one basic block
            double lnb[DM] = {}, tnb[DM] = {}, fh[DM] = {};
            double lo_own = 0.0;
            if (rd) {
                Tc[s] = Tr[k];
                Ac[s] = ab2(Or[k], Tc[s]);
                lo_own = LOr[k];
#pragma unroll
                for (int q = 0; q < DM; ++q) {
                    const size_t r = (size_t)es[q].y * L + k;
                    lnb[q] = LOb[r];
                    tnb[q] = Tb[r];
                    fh[q] = adfh[(size_t)(es[q].x >> 1) * L + k];
                }
            }
            LOc[s] = wet ? lo_own : 0.0;
            // a1-a3: max/min of (LO, T) over the vertices of the node's triangles that are
            // wet at k, and the -+1e3 pad when one of them is dry at k
            bool act = false, inact = false;
#pragma unroll
            for (int q = 0; q < DM; ++q) {
                const bool in = in_bytes(es[q].w, 0, k);
                act = act || (q < cnt.y && in);
                inact = inact || (q < cnt.y && !in);
            }
            double mx = act ? fmax(LOc[s], Tc[s]) : -INFINITY;
            double mn = act ? fmin(LOc[s], Tc[s]) : INFINITY;
#pragma unroll
            for (int q = 0; q < DM; ++q) {
                const bool inc = q < cnt.x && (in_bytes(es[q].z, 0, k) || in_bytes(es[q].z, 16, k));
                mx = inc ? fmax(mx, fmax(lnb[q], tnb[q])) : mx;
                mn = inc ? fmin(mn, fmin(lnb[q], tnb[q])) : mn;
            }
            mx = inact ? fmax(mx, -FCT_BIG) : mx;
            mn = inact ? fmin(mn, FCT_BIG) : mn;
            tvx[s] = rd ? mx : -INFINITY;
            tvn[s] = rd ? mn : INFINITY;
            // b1: P+/P- from the incident edges in reference order
#pragma unroll
            for (int q = 0; q < DM; ++q) {
                const double f = (es[q].x & 1) ? -fh[q] : fh[q];
                const bool use = wet && q < cnt.x;
                ap[s] = ap[s] + (use ? fmax(f, 0.0) : 0.0);
                am[s] = am[s] + (use ? fmin(f, 0.0) : 0.0);
            }
        }
        double tvxm[NS], tvxp[NS], tvnm[NS], tvnp[NS];
        col_prev<NS>(tvx, tvxm, col_bcast0<NS>(tvx));
        col_next<NS>(tvx, tvxp, 0.0);
        col_prev<NS>(tvn, tvnm, col_bcast0<NS>(tvn));
        col_next<NS>(tvn, tvnp, 0.0);
        double fvc[NS], fvb[NS];
        adfv_compute<NS>(a, nr, Tc, Ac, Ws, ars, Zcs, zbs, fvc);
        col_next<NS>(fvc, fvb, 0.0);
        // for pass 3, which reads it instead of computing it again
        const auto Fv = G(a.adfv + jo);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int k = 64 * s + lane_id();
            if (k <= nr.y && k < L) Fv[b0 + k] = fvc[s];
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int k = 64 * s + lane_id();
            if (!in_layer(nr, k)) continue;
            const size_t i = b0 + k;
            const bool use_own = (k == nr.x) || (k == nr.y - 1);
            const double cmax = use_own ? tvx[s] : fmax(fmax(tvxm[s], tvx[s]), tvxp[s]);
            const double cmin = use_own ? tvn[s] : fmin(fmin(tvnm[s], tvn[s]), tvnp[s]);
            const double ttf_max = cmax - LOc[s];
            const double ttf_min = cmin - LOc[s];
            const double pos_v = fmax(fvc[s], 0.0) + fmax(-fvb[s], 0.0);
            const double neg_v = fmin(fvc[s], 0.0) + fmin(-fvb[s], 0.0);
            const double fplus = pos_v + ap[s];
            const double fminus = neg_v + am[s];
            const double av = avs[s], hn = hns[s];
            const double safe_a = av > 0.0 ? av : 1.0, safe_hn = hn > 0.0 ? hn : 1.0;
            const double flux_pos = fplus * a.dt / safe_a / safe_hn + FCT_FLUX_EPS;
            const double flux_neg = fminus * a.dt / safe_a / safe_hn - FCT_FLUX_EPS;
            const bool valid = av > 0.0 && hn > 0.0;
            // fmin(1, x / y) is 1 wherever x / y >= 1 (rounding is monotone and 1 is a
            // double), which the signs and one comparison decide: only lanes where the
            // limiter may bind divide, and a wavefront with none skips the divisions
            const bool bp = valid && !(flux_pos > 0.0 && ttf_max >= flux_pos);
            const bool bm = valid && !(flux_neg < 0.0 && ttf_min <= flux_neg);
            double fp = 1.0, fm = 1.0;
            if (bp) fp = fmin(1.0, ttf_max / flux_pos);
            if (bm) fm = fmin(1.0, ttf_min / flux_neg);
            // (plus, minus) factor pairs, interleaved: pass 3 reads both with one load
            stv(reinterpret_cast<double2*>(a.pf), jo + i, make_double2(fp, fm));
        }
    }
}

template <int NS, int DM>
__device__ __forceinline__ void pass2_task(const FusedArgs& a, int n, int j) {
    const int4 hd = ldc(a.tab.hdr, n);
    const int2 nr = make_int2(hd.x, hd.y), cnt = make_int2(hd.z & 0xff, hd.w);
    if (NS == 1 || nr.y < 64) pass2_body<1, DM>(a, n, j, nr, cnt);
    else pass2_body<NS, DM>(a, n, j, nr, cnt);
}

// ---------------------------------------------------------------------------
// pass 3 of node n, tracer j: limited fluxes, divergence, reconstruction
// ---------------------------------------------------------------------------
template <int NS, int DM>
__device__ __forceinline__ void pass3_body(const FusedArgs& a, int n, int j, int2 nr, int2 cnt) {
    const DevMesh& m = a.m;
    const int L = m.L;
    const auto T = G(a.T[j]);
    const size_t jo = (size_t)j * m.N * L;
    const double2* pm = reinterpret_cast<const double2*>(a.pf) + jo;   // (plus, minus) pairs
    const auto adfh = G(a.adfh + (size_t)j * m.D * L);
    const size_t b0 = (size_t)n * L;
    const int kneed = nr.y + 1 < L ? nr.y + 1 : L;
    int4 es[DM];
#pragma unroll
    for (int q = 0; q < DM; ++q) es[q] = ldc(a.tab.edge, (size_t)n * DM + q);
    // adf_v at the node's interfaces (k <= nzmax) as pass 2 computed it
    const auto Fv = G(a.adfv + jo);
    double Tc[NS], fvc[NS], Pc[NS], Mc[NS], acc[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        Tc[s] = 0.0; fvc[s] = 0.0; Pc[s] = 1.0; Mc[s] = 1.0; acc[s] = 0.0;
        if (64 * s >= kneed) continue;
        const int k = 64 * s + lane_id();
        const bool wet = in_layer(nr, k);
        // T on every level of the slot (one load for the wet levels and the dry ones
        // below, which are copied)
        if (k < L) Tc[s] = T[b0 + k];
        if (k < kneed) fvc[s] = Fv[b0 + k];
        double fh[DM] = {}, po[DM] = {}, mo[DM] = {};
        if (wet) {
#pragma unroll
            for (int q = 0; q < DM; ++q) {
                fh[q] = adfh[(size_t)(es[q].x >> 1) * L + k];
                const double2 f = ldv(pm, (size_t)es[q].y * L + k);
                po[q] = f.x;
                mo[q] = f.y;
            }
        }
        if (wet) {
            const double2 f = ldv(pm, b0 + k);
            Pc[s] = f.x;
            Mc[s] = f.y;
        }
#pragma unroll
        for (int q = 0; q < DM; ++q) {
            const double f = fh[q];
            const bool side = es[q].x & 1;
            const double ae = side ? (f >= 0.0 ? fmin(po[q], Mc[s]) : fmin(mo[q], Pc[s]))
                                   : (f >= 0.0 ? fmin(Pc[s], mo[q]) : fmin(Mc[s], po[q]));
            const double fl = ae * f;
            const bool use = wet && q < cnt.x && f != 0.0;
            acc[s] = acc[s] + (use ? (side ? -fl : fl) : 0.0);
        }
    }
    double Pm[NS], Mm[NS];
    col_prev<NS>(Pc, Pm, col_bcast0<NS>(Pc));
    col_prev<NS>(Mc, Mm, col_bcast0<NS>(Mc));
    double vl[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int k = 64 * s + lane_id();
        vl[s] = 0.0;
        if (in_iface(nr, k) && k < L) {
            const double f = fvc[s];
            double ae;
            if (k == nr.x) ae = f >= 0.0 ? Pc[s] : Mc[s];
            else ae = f >= 0.0 ? fmin(Mm[s], Pc[s]) : fmin(Pm[s], Mc[s]);
            vl[s] = ae * f;
        }
    }
    double vl1[NS];
    col_next<NS>(vl, vl1, 0.0);
    const auto LO = G(a.LO + jo);
    const auto Tn = G(a.Tnew[j]);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int k = 64 * s + lane_id();
        if (k >= L) continue;
        const size_t i = b0 + k;
        if (!in_layer(nr, k)) {
            const double Ti = 64 * s < kneed ? Tc[s] : T[i];
            Tn[i] = Ti + 0.0;
            continue;
        }
        const double av = G(m.areasvol)[i], hn = G(a.hnn)[i], hnd = G(a.hnode)[i];
        const double safe_a = av > 0.0 ? av : 1.0, safe_hn = hn > 0.0 ? hn : 1.0;
        const double Ti = Tc[s];
        const double lo_trans = -Ti * hnd + LO[i] * hn;
        const double vdiv = av > 0.0 ? (vl[s] - vl1[s]) * a.dt / safe_a : 0.0;
        const double dttf_v = lo_trans + vdiv;
        const double dttf_h = acc[s] * a.dt / safe_a;
        double del = dttf_h + dttf_v;
        del = del + Ti * (hnd - hn);
        const double dT = hn > 0.0 ? del / safe_hn : 0.0;
        Tn[i] = Ti + dT;
    }
    // levels past the slots (a shallow column taking the one-slot body): dry copies
    for (int k = 64 * NS + lane_id(); k < L; k += 64) Tn[b0 + k] = T[b0 + k] + 0.0;
}

// Pass 3 of a shallow node (one 64-level slot) for its KT tracers in one wavefront:
// areasvol, hnode_new and hnode are loaded once; per tracer the arithmetic is
// pass3_body's.
template <int DM, int KT>
__device__ __forceinline__ void pass3_shallow_all(const FusedArgs& a, int n, int2 nr, int2 cnt, int j0) {
    constexpr int NS = 1;
    const DevMesh& m = a.m;
    const int L = m.L;
    int4 es[DM];
#pragma unroll
    for (int q = 0; q < DM; ++q) es[q] = ldc(a.tab.edge, (size_t)n * DM + q);
    const int kl = lane_id();
    const bool wetl = in_layer(nr, kl);
    double av_s = 0.0, hn_s = 0.0, hnd_s = 0.0;
    if (wetl) {
        av_s = G(m.areasvol)[(size_t)n * L + kl];
        hn_s = G(a.hnn)[(size_t)n * L + kl];
        hnd_s = G(a.hnode)[(size_t)n * L + kl];
    }
#pragma unroll
    for (int jj = 0; jj < KT; ++jj) {
        const int j = j0 + jj;   // tracer slot
        const auto T = G(a.T[j]);
        const size_t jo = (size_t)j * m.N * L;
        const double2* pm = reinterpret_cast<const double2*>(a.pf) + jo;   // (plus, minus) pairs
        const auto adfh = G(a.adfh + (size_t)j * m.D * L);
        const size_t b0 = (size_t)n * L;
        const int kneed = nr.y + 1 < L ? nr.y + 1 : L;
        // adf_v at the node's interfaces (k <= nzmax) as pass 2 computed it
        const auto Fv = G(a.adfv + jo);
        double Tc[NS], fvc[NS], Pc[NS], Mc[NS], acc[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            Tc[s] = 0.0; fvc[s] = 0.0; Pc[s] = 1.0; Mc[s] = 1.0; acc[s] = 0.0;
            if (64 * s >= kneed) continue;
            const int k = 64 * s + lane_id();
            const bool wet = in_layer(nr, k);
            // T on every level of the slot (one load for the wet levels and the dry ones
            // below, which are copied)
            if (k < L) Tc[s] = T[b0 + k];
            if (k < kneed) fvc[s] = Fv[b0 + k];
            double fh[DM] = {}, po[DM] = {}, mo[DM] = {};
            if (wet) {
#pragma unroll
                for (int q = 0; q < DM; ++q) {
                    fh[q] = adfh[(size_t)(es[q].x >> 1) * L + k];
                    const double2 f = ldv(pm, (size_t)es[q].y * L + k);
                    po[q] = f.x;
                    mo[q] = f.y;
                }
            }
            if (wet) {
                const double2 f = ldv(pm, b0 + k);
                Pc[s] = f.x;
                Mc[s] = f.y;
            }
#pragma unroll
            for (int q = 0; q < DM; ++q) {
                const double f = fh[q];
                const bool side = es[q].x & 1;
                const double ae = side ? (f >= 0.0 ? fmin(po[q], Mc[s]) : fmin(mo[q], Pc[s]))
                                       : (f >= 0.0 ? fmin(Pc[s], mo[q]) : fmin(Mc[s], po[q]));
                const double fl = ae * f;
                const bool use = wet && q < cnt.x && f != 0.0;
                acc[s] = acc[s] + (use ? (side ? -fl : fl) : 0.0);
            }
        }
        double Pm[NS], Mm[NS];
        col_prev<NS>(Pc, Pm, col_bcast0<NS>(Pc));
        col_prev<NS>(Mc, Mm, col_bcast0<NS>(Mc));
        double vl[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int k = 64 * s + lane_id();
            vl[s] = 0.0;
            if (in_iface(nr, k) && k < L) {
                const double f = fvc[s];
                double ae;
                if (k == nr.x) ae = f >= 0.0 ? Pc[s] : Mc[s];
                else ae = f >= 0.0 ? fmin(Mm[s], Pc[s]) : fmin(Pm[s], Mc[s]);
                vl[s] = ae * f;
            }
        }
        double vl1[NS];
        col_next<NS>(vl, vl1, 0.0);
        const auto LO = G(a.LO + jo);
        const auto Tn = G(a.Tnew[j]);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int k = 64 * s + lane_id();
            if (k >= L) continue;
            const size_t i = b0 + k;
            if (!in_layer(nr, k)) {
                const double Ti = 64 * s < kneed ? Tc[s] : T[i];
                Tn[i] = Ti + 0.0;
                continue;
            }
            const double av = av_s, hn = hn_s, hnd = hnd_s;
            const double safe_a = av > 0.0 ? av : 1.0, safe_hn = hn > 0.0 ? hn : 1.0;
            const double Ti = Tc[s];
            const double lo_trans = -Ti * hnd + LO[i] * hn;
            const double vdiv = av > 0.0 ? (vl[s] - vl1[s]) * a.dt / safe_a : 0.0;
            const double dttf_v = lo_trans + vdiv;
            const double dttf_h = acc[s] * a.dt / safe_a;
            double del = dttf_h + dttf_v;
            del = del + Ti * (hnd - hn);
            const double dT = hn > 0.0 ? del / safe_hn : 0.0;
            Tn[i] = Ti + dT;
        }
        // levels past the slots (a shallow column taking the one-slot body): dry copies
        for (int k = 64 * NS + lane_id(); k < L; k += 64) Tn[b0 + k] = T[b0 + k] + 0.0;
    }
}

template <int NS, int DM>
__device__ __forceinline__ void pass3_task(const FusedArgs& a, int n, int j) {
    const int4 hd = ldc(a.tab.hdr, n);
    const int2 nr = make_int2(hd.x, hd.y), cnt = make_int2(hd.z & 0xff, hd.w);
    if (NS == 1 || nr.y < 64) pass3_body<1, DM>(a, n, j, nr, cnt);
    else pass3_body<NS, DM>(a, n, j, nr, cnt);
}

// ---------------------------------------------------------------------------
// upwind-only step of node n, tracer j (advect_one, tracer_adv.py:154-165): upwind
// fluxes of ttfAB (adv_flux_hor :96-105, adv_flux_ver :108-120), flux2dtracer
// (:123-139) and ale_reconstruct (:142-151); every lane of T_new is written
// ---------------------------------------------------------------------------
template <int NS, int DM>
__device__ __forceinline__ void upwind_task(const FusedArgs& a, int n, int j) {
    const DevMesh& m = a.m;
    const int L = m.L;
    const int4 hd = ldc(a.tab.hdr, n);
    const int2 nr = make_int2(hd.x, hd.y), cnt = make_int2(hd.z & 0xff, hd.w);
    const auto T = G(a.T[j]);
    const auto To = G(a.Told[j]);
    const auto Tn = G(a.Tnew[j]);
    const size_t b0 = (size_t)n * L;
    int4 es[DM];
#pragma unroll
    for (int q = 0; q < DM; ++q) es[q] = ldc(a.tab.edge, (size_t)n * DM + q);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int k = 64 * s + lane_id();
        if (64 * s >= L) break;
        const bool wet = in_layer(nr, k);
        const size_t i = b0 + k;
        double Ti = 0.0;
        if (k < L) Ti = T[i];
        if (!wet) {
            if (k < L) Tn[i] = Ti + 0.0;
            continue;
        }
        const double A = ab2(To[i], Ti);
        const size_t im = k > 0 ? i - 1 : i;
        const double Am = ab2(To[im], T[im]);
        const bool below = in_layer(nr, k + 1);
        const double Ap = below ? ab2(To[i + 1], T[i + 1]) : 0.0;
        double vf[DM], A2[DM];
#pragma unroll
        for (int q = 0; q < DM; ++q) {
            vf[q] = G(a.vflux)[(size_t)(es[q].x >> 1) * L + k];
            const size_t r = (size_t)es[q].y * L + k;
            A2[q] = ab2(To[r], T[r]);
        }
        double raw = 0.0;
#pragma unroll
        for (int q = 0; q < DM; ++q) {
            const double v = vf[q];
            const bool side = es[q].x & 1;
            const double T1 = side ? A2[q] : A;
            const double T2 = side ? A : A2[q];
            const double av = fabs(v);
            const double f = -0.5 * (T1 * (v + av) + T2 * (v - av));
            const bool use = q < cnt.x && v != 0.0;
            raw = raw + (use ? (side ? -f : f) : 0.0);
        }
        double fv0, fv1 = 0.0;
        {
            const double wk = G(a.w_e)[i], aw = fabs(wk);
            fv0 = (-0.5 * (A * (wk + aw) + Am * (wk - aw))) * G(m.area)[i];
        }
        if (below) {
            const double wk = G(a.w_e)[i + 1], aw = fabs(wk);
            fv1 = (-0.5 * (Ap * (wk + aw) + A * (wk - aw))) * G(m.area)[i + 1];
        }
        const double av = G(m.areasvol)[i], hn = G(a.hnn)[i], hnd = G(a.hnode)[i];
        const double safe_a = av > 0.0 ? av : 1.0, safe_hn = hn > 0.0 ? hn : 1.0;
        const double dttf_v = (fv0 - fv1) * a.dt / safe_a;
        const double dttf_h = raw * a.dt / safe_a;
        double del = dttf_h + dttf_v;
        del = del + Ti * (hnd - hn);
        const double dT = hn > 0.0 ? del / safe_hn : 0.0;
        Tn[i] = Ti + dT;
    }
}

// ---------------------------------------------------------------------------
// kernels: single instance (descriptor by value) and batched (descriptor array)
// ---------------------------------------------------------------------------
// Occupancy targets (wavefronts per SIMD) per kernel: the most the compiler reaches
// without spilling to scratch (node degree 8 tables need more registers); the passes
// are latency-bound, so more resident wavefronts hide more of each one's memory waits.
#define FCT_OCC(n) __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(n)))

template <int NS, int EPW>
__global__ void FCT_OCC(8) kf_vflux(FusedArgs a) {
    int w;
    if (single_wave((a.m.D + EPW - 1) / EPW, a.remap, w)) vflux_group<NS, EPW>(a, w);
}
template <int NS, int EPW>
__global__ void FCT_OCC(8) kb_vflux(const FusedArgs* __restrict__ d,
                                                const int2* __restrict__ blk, int remap) {
    int inst;
    const int w = batch_wave(blk, remap, inst);
    const FusedArgs& a = d[inst];
    if (w < (a.m.D + EPW - 1) / EPW) vflux_group<NS, EPW>(a, w);
}

template <int NS, int EPW>
__global__ void FCT_OCC(8) kf_adfh(FusedArgs a) {
    int w;
    if (!single_wave((int64_t)((a.m.D + EPW - 1) / EPW) * a.K, a.remap, w)) return;
    const int g = w / a.K;
    adfh_group<NS, EPW>(a, g, w - g * a.K);
}
template <int NS, int EPW>
__global__ void FCT_OCC(8) kb_adfh(const FusedArgs* __restrict__ d,
                                               const int2* __restrict__ blk, int remap) {
    int inst;
    const int w = batch_wave(blk, remap, inst);
    const FusedArgs& a = d[inst];
    if (w >= ((a.m.D + EPW - 1) / EPW) * a.K) return;
    const int g = w / a.K;
    adfh_group<NS, EPW>(a, g, w - g * a.K);
}

// adf_h over edge patches; KT tracers from slot j0 of the launch's tracer chunk
// at most 128 registers per thread
template <int KT, int KC, int TPE>
__global__ void __launch_bounds__(FCT_EP * TPE) __attribute__((amdgpu_waves_per_eu(4)))
ke_adfh(FusedArgs a, int j0) {
    adfh_patch<KT, KC, TPE>(a, (int)block_tile(a.remap), j0);
}
// batched form: per workgroup (instance, patch | first tracer << 24)
template <int KT, int KC, int TPE>
__global__ void __launch_bounds__(FCT_EP * TPE) __attribute__((amdgpu_waves_per_eu(4)))
kbe_adfh(const FusedArgs* __restrict__ d, const int2* __restrict__ blk, int remap) {
    const int2 e = blk[block_tile(remap)];
    const int inst = __builtin_amdgcn_readfirstlane(e.x);
    const int pj = __builtin_amdgcn_readfirstlane(e.y);
    adfh_patch<KT, KC, TPE>(d[inst], pj & 0xffffff, pj >> 24);
}

#define FCT_NODE_KERNELS(PASS, OCC)                                                     \
    template <int NS, int DM>                                                           \
    __global__ void FCT_OCC(DM == 6 ? OCC : 4) kf_##PASS(FusedArgs a) {                 \
        int w;                                                                          \
        if (!single_wave((int64_t)a.m.N * a.K, a.remap, w)) return;                       \
        const int n = w / a.K;                                                          \
        PASS##_task<NS, DM>(a, n, w - n * a.K);                                         \
    }                                                                                   \
    template <int NS, int DM>                                                           \
    __global__ void FCT_OCC(DM == 6 ? OCC : 4) kb_##PASS(const FusedArgs* __restrict__ d, \
                                                     const int2* __restrict__ blk,      \
                                                     int remap) {                         \
        int inst;                                                                       \
        const int w = batch_wave(blk, remap, inst);                                       \
        const FusedArgs& a = d[inst];                                                   \
        if (w >= a.m.N * a.K) return;                                                   \
        const int n = w / a.K;                                                          \
        PASS##_task<NS, DM>(a, n, w - n * a.K);                                         \
    }

FCT_NODE_KERNELS(upwind, 5)
FCT_NODE_KERNELS(pass1, 5)
FCT_NODE_KERNELS(pass2, 5)
FCT_NODE_KERNELS(pass3, 6)
#undef FCT_NODE_KERNELS

// Passes 2 and 3 over `count` nodes of the node list from `first`, all of whose wet
// layers fit NSB slots: shallow columns (NSB = 1) run with fewer registers per
// wavefront and so more wavefronts per SIMD than the two-slot body needs.
#define FCT_LIST_KERNELS(PASS, OCC1, OCC2)                                                  template <int NSB, int DM>                                                              __global__ void FCT_OCC(DM == 6 ? (NSB == 1 ? OCC1 : OCC2) : 4)                              kl_##PASS(FusedArgs a, int first, int count) {                                          int w;                                                                                  if (!single_wave((int64_t)count * a.K, a.remap, w)) return;                              const int t = w / a.K;                                                                  const int n = ldc(a.tab.nlist, (size_t)first + t);                                      const int4 hd = ldc(a.tab.hdr, n);                                                      PASS##_body<NSB, DM>(a, n, w - t * a.K, make_int2(hd.x, hd.y),                                               make_int2(hd.z & 0xff, hd.w));                                 }

FCT_LIST_KERNELS(pass2, 7, 4)
FCT_LIST_KERNELS(pass3, 8, 6)
#undef FCT_LIST_KERNELS

// batched form: the block table maps each workgroup to (instance, first wavefront of the
// instance's shallow (deep = 0) or deep (deep = 1) node list)
#define FCT_BATCH_LIST_KERNEL(PASS, OCC1, OCC2, CALL)                                      \
    template <int NSB, int DM>                                                            \
    __global__ void FCT_OCC(DM == 6 ? (NSB == 1 ? OCC1 : OCC2) : 4)                        \
        kbl_##PASS(const FusedArgs* __restrict__ d, const int2* __restrict__ blk, int remap, \
                   int deep) {                                                            \
        int inst;                                                                         \
        const int w = batch_wave(blk, remap, inst);                                       \
        const FusedArgs& a = d[inst];                                                     \
        const int ns = a.tab.nshallow;                                                    \
        const int first = deep ? ns : 0, count = deep ? a.m.N - ns : ns;                  \
        if (w >= count * a.K) return;                                                     \
        const int t = w / a.K;                                                            \
        const int n = ldc(a.tab.nlist, (size_t)first + t);                                \
        const int4 hd = ldc(a.tab.hdr, n);                                                \
        CALL;                                                                             \
    }

FCT_BATCH_LIST_KERNEL(pass1, 5, 5, (pass1_body<NSB, DM>(a, n, w - t * a.K, hd)))
FCT_BATCH_LIST_KERNEL(pass2, 7, 4, (pass2_body<NSB, DM>(a, n, w - t * a.K, make_int2(hd.x, hd.y),
                                                        make_int2(hd.z & 0xff, hd.w))))
FCT_BATCH_LIST_KERNEL(pass3, 8, 6, (pass3_body<NSB, DM>(a, n, w - t * a.K, make_int2(hd.x, hd.y),
                                                        make_int2(hd.z & 0xff, hd.w))))
#undef FCT_BATCH_LIST_KERNEL

// one wavefront per shallow node for KT tracers from slot j0
template <int DM, int KT>
__global__ void FCT_OCC(DM == 6 ? 6 : 4) kl_pass2_all(FusedArgs a, int count, int j0) {
    int w;
    if (!single_wave((int64_t)count, a.remap, w)) return;
    const int n = ldc(a.tab.nlist, (size_t)w);
    const int4 hd = ldc(a.tab.hdr, n);
    pass2_shallow_all<DM, KT>(a, n, make_int2(hd.x, hd.y), make_int2(hd.z & 0xff, hd.w), j0);
}

template <int DM, int KT>
__global__ void FCT_OCC(DM == 6 ? 7 : 4) kl_pass3_all(FusedArgs a, int count, int j0) {
    int w;
    if (!single_wave((int64_t)count, a.remap, w)) return;
    const int n = ldc(a.tab.nlist, (size_t)w);
    const int4 hd = ldc(a.tab.hdr, n);
    pass3_shallow_all<DM, KT>(a, n, make_int2(hd.x, hd.y), make_int2(hd.z & 0xff, hd.w), j0);
}

template <int DM, int KT>
__global__ void FCT_OCC(DM == 6 ? 5 : 4) kl_pass1_all(FusedArgs a, int count, int j0) {
    int w;
    if (!single_wave((int64_t)count, a.remap, w)) return;
    const int n = ldc(a.tab.nlist, (size_t)w);
    pass1_shallow_all<DM, KT>(a, n, ldc(a.tab.hdr, n), j0);
}

// batched multi-tracer passes: waves over each instance's shallow node list; the block
// table's instance field carries the first tracer slot in its high 16 bits
template <int DM, int KT>
__global__ void FCT_OCC(DM == 6 ? 5 : 4) kbp_pass1(const FusedArgs* __restrict__ d,
                                                     const int2* __restrict__ blk, int remap) {
    int ij;
    const int w = batch_wave(blk, remap, ij);
    const FusedArgs& a = d[ij & 0xffff];
    if (w >= a.tab.nshallow) return;
    const int n = ldc(a.tab.nlist, (size_t)w);
    pass1_shallow_all<DM, KT>(a, n, ldc(a.tab.hdr, n), ij >> 16);
}
template <int DM, int KT>
__global__ void FCT_OCC(DM == 6 ? 6 : 4) kbp_pass2(const FusedArgs* __restrict__ d,
                                                     const int2* __restrict__ blk, int remap) {
    int ij;
    const int w = batch_wave(blk, remap, ij);
    const FusedArgs& a = d[ij & 0xffff];
    if (w >= a.tab.nshallow) return;
    const int n = ldc(a.tab.nlist, (size_t)w);
    const int4 hd = ldc(a.tab.hdr, n);
    pass2_shallow_all<DM, KT>(a, n, make_int2(hd.x, hd.y), make_int2(hd.z & 0xff, hd.w), ij >> 16);
}
template <int DM, int KT>
__global__ void FCT_OCC(DM == 6 ? 7 : 4) kbp_pass3(const FusedArgs* __restrict__ d,
                                                     const int2* __restrict__ blk, int remap) {
    int ij;
    const int w = batch_wave(blk, remap, ij);
    const FusedArgs& a = d[ij & 0xffff];
    if (w >= a.tab.nshallow) return;
    const int n = ldc(a.tab.nlist, (size_t)w);
    const int4 hd = ldc(a.tab.hdr, n);
    pass3_shallow_all<DM, KT>(a, n, make_int2(hd.x, hd.y), make_int2(hd.z & 0xff, hd.w), ij >> 16);
}

template <int NSB, int DM>
__global__ void FCT_OCC(DM == 6 ? 8 : 4) kl_pass1(FusedArgs a, int first,
                                                                   int count) {
    int w;
    if (!single_wave((int64_t)count * a.K, a.remap, w)) return;
    const int t = w / a.K;
    const int n = ldc(a.tab.nlist, (size_t)first + t);
    pass1_body<NSB, DM>(a, n, w - t * a.K, ldc(a.tab.hdr, n));
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static inline unsigned blocks_for(int64_t waves, int wpb) {
    return (unsigned)((waves + wpb - 1) / wpb);
}

// node pass `stage` (1, 3 or 4) over the first `count` nodes of the node list with the
// tracers taken two per wavefront (and a last one alone when K is odd)
template <int DM>
static void launch_shallow_multi(const FusedArgs& fa, const FusedTuning& t, hipStream_t s,
                                 int stage, int count) {
    const dim3 blk(64 * t.wpb), g(blocks_for(count, t.wpb));
    for (int j0 = 0; j0 < fa.K; j0 += 2) {
        const bool two = fa.K - j0 >= 2;
        if (stage == 1) {
            if (two) hipLaunchKernelGGL((kl_pass1_all<DM, 2>), g, blk, 0, s, fa, count, j0);
            else hipLaunchKernelGGL((kl_pass1_all<DM, 1>), g, blk, 0, s, fa, count, j0);
        } else if (stage == 3) {
            if (two) hipLaunchKernelGGL((kl_pass2_all<DM, 2>), g, blk, 0, s, fa, count, j0);
            else hipLaunchKernelGGL((kl_pass2_all<DM, 1>), g, blk, 0, s, fa, count, j0);
        } else {
            if (two) hipLaunchKernelGGL((kl_pass3_all<DM, 2>), g, blk, 0, s, fa, count, j0);
            else hipLaunchKernelGGL((kl_pass3_all<DM, 1>), g, blk, 0, s, fa, count, j0);
        }
    }
}

template <int NS, int DM>
static void launch_single(const FusedArgs& fa, const FusedTuning& t, hipStream_t s, int stage) {
    const DevMesh& m = fa.m;
    const dim3 blk(64 * t.wpb);
    const unsigned gn = blocks_for(node_waves(m.N, fa.K), t.wpb);
    // several tracers share the tracer-independent loads of a shallow node in one
    // wavefront (two at a time)
    const bool multi = DM == 6 && fa.K >= 2 && t.pair && fa.tab.nlist;
    switch (stage) {
        case 0:
            if (t.vflux_epw == 2)
                hipLaunchKernelGGL((kf_vflux<NS, 2>), dim3(blocks_for(vflux_waves(m.D, t), t.wpb)), blk, 0, s, fa);
            else
                hipLaunchKernelGGL((kf_vflux<NS, 4>), dim3(blocks_for(vflux_waves(m.D, t), t.wpb)), blk, 0, s, fa);
            break;
        case 1:
            if (NS == 1 && multi) {
                // nl <= 64: every column is shallow
                launch_shallow_multi<DM>(fa, t, s, 1, fa.tab.nshallow);
            } else if (NS == 1 && fa.tab.nlist) {
                // nl <= 64 in the list's (visit) order
                hipLaunchKernelGGL((kl_pass1<1, DM>), dim3(gn), blk, 0, s, fa, 0, m.N);
            } else if (NS == 2 && fa.tab.nlist) {
                const int ns = fa.tab.nshallow, nd = m.N - fa.tab.nshallow;
                if (ns && multi)
                    launch_shallow_multi<DM>(fa, t, s, 1, ns);
                else if (ns)
                    hipLaunchKernelGGL((kl_pass1<1, DM>), dim3(blocks_for(node_waves(ns, fa.K), t.wpb)),
                                       blk, 0, s, fa, 0, ns);
                if (nd) hipLaunchKernelGGL((kl_pass1<2, DM>), dim3(blocks_for(node_waves(nd, fa.K), t.wpb)),
                                           blk, 0, s, fa, ns, nd);
            } else {
                hipLaunchKernelGGL((kf_pass1<NS, DM>), dim3(gn), blk, 0, s, fa);
            }
            break;
        case 2:
            if (t.adfh_kc && fa.tab.ep.np) {
                const dim3 g(fa.tab.ep.np);
                for (int j = 0; j < fa.K; j += 2) {
                    const bool two = fa.K - j >= 2;
#define FCT_KE(KC, TPE)                                                                         \
    do {                                                                                        \
        if (two) hipLaunchKernelGGL((ke_adfh<2, KC, TPE>), g, dim3(FCT_EP * TPE), 0, s, fa, j); \
        else hipLaunchKernelGGL((ke_adfh<1, KC, TPE>), g, dim3(FCT_EP * TPE), 0, s, fa, j);     \
    } while (0)
                    if (t.adfh_kc == 4) {
                        if (t.adfh_tpe == 1) FCT_KE(4, 1);
                        else if (t.adfh_tpe == 2) FCT_KE(4, 2);
                        else FCT_KE(4, 4);
                    } else {
                        if (t.adfh_tpe == 1) FCT_KE(8, 1);
                        else if (t.adfh_tpe == 2) FCT_KE(8, 2);
                        else FCT_KE(8, 4);
                    }
#undef FCT_KE
                }
            } else if (t.adfh_epw == 1)
                hipLaunchKernelGGL((kf_adfh<NS, 1>), dim3(blocks_for(adfh_waves(m.D, fa.K, t), t.wpb)), blk, 0, s, fa);
            else
                hipLaunchKernelGGL((kf_adfh<NS, 2>), dim3(blocks_for(adfh_waves(m.D, fa.K, t), t.wpb)), blk, 0, s, fa);
            break;
        case 3:
        case 4: {
            const NodeTab& tb = fa.tab;
            if (NS == 1 && multi) {
                launch_shallow_multi<DM>(fa, t, s, stage, tb.nshallow);
            } else if (NS == 1 && tb.nlist) {
                if (stage == 3) hipLaunchKernelGGL((kl_pass2<1, DM>), dim3(gn), blk, 0, s, fa, 0, m.N);
                else hipLaunchKernelGGL((kl_pass3<1, DM>), dim3(gn), blk, 0, s, fa, 0, m.N);
            } else if (NS == 2 && tb.nlist) {
                const int ns = tb.nshallow, nd = m.N - tb.nshallow;
                const unsigned g1 = blocks_for(node_waves(ns, fa.K), t.wpb);
                const unsigned g2 = blocks_for(node_waves(nd, fa.K), t.wpb);
                if (stage == 3) {
                    if (ns && multi)
                        launch_shallow_multi<DM>(fa, t, s, 3, ns);
                    else if (ns)
                        hipLaunchKernelGGL((kl_pass2<1, DM>), dim3(g1), blk, 0, s, fa, 0, ns);
                    if (nd) hipLaunchKernelGGL((kl_pass2<2, DM>), dim3(g2), blk, 0, s, fa, ns, nd);
                } else {
                    if (ns && multi)
                        launch_shallow_multi<DM>(fa, t, s, 4, ns);
                    else if (ns)
                        hipLaunchKernelGGL((kl_pass3<1, DM>), dim3(g1), blk, 0, s, fa, 0, ns);
                    if (nd) hipLaunchKernelGGL((kl_pass3<2, DM>), dim3(g2), blk, 0, s, fa, ns, nd);
                }
            } else if (stage == 3) {
                hipLaunchKernelGGL((kf_pass2<NS, DM>), dim3(gn), blk, 0, s, fa);
            } else {
                hipLaunchKernelGGL((kf_pass3<NS, DM>), dim3(gn), blk, 0, s, fa);
            }
            break;
        }
    }
}

template <int NS, int DM>
static int launch_fused_t(FusedArgs& fa, const fct_step_args* a, int kw, const FusedTuning& t,
                          hipStream_t s) {
    fa.K = 1;
    if (fa.m.D) launch_single<NS, DM>(fa, t, s, 0);
    int rc = check_launch("kf_vflux");
    if (rc) return rc;
    for (int j0 = 0; j0 < a->ntracers; j0 += kw) {
        const int K = (a->ntracers - j0) < kw ? (a->ntracers - j0) : kw;
        fa.K = K;
        for (int j = 0; j < FCT_KCHUNK; ++j) {
            const int jj = j0 + (j < K ? j : 0);
            fa.T[j] = a->T[jj]; fa.Told[j] = a->T_old[jj]; fa.Tnew[j] = a->T_new[jj];
        }
        for (int stage = 1; stage <= 4; ++stage) launch_single<NS, DM>(fa, t, s, stage);
        if ((rc = check_launch("fused passes"))) return rc;
    }
    return FCT_OK;
}

static bool tuning_ok(const FusedTuning& t) {
    return (t.wpb == 1 || t.wpb == 2 || t.wpb == 4) && (t.vflux_epw == 2 || t.vflux_epw == 4) &&
           (t.adfh_epw == 1 || t.adfh_epw == 2) && t.xcd >= 1 &&
           (t.adfh_kc == 0 || t.adfh_kc == 4 || t.adfh_kc == 8) &&
           (t.adfh_tpe == 1 || t.adfh_tpe == 2 || t.adfh_tpe == 4);
}

static void base_args(FusedArgs& fa, const DevMesh& m, const NodeTab& tab, const int2* ekr,
                      const fct_step_args* a, const FusedTuning& tune, double* vflux, double* LO,
                      double* gnode, double* adfh, double* pf, double* adfv) {
    fa.m = m; fa.tab = tab; fa.vflux = vflux; fa.ekr = ekr; fa.uv = a->uv; fa.helem = a->helem;
    fa.w_e = a->w_e; fa.hnode = a->hnode; fa.hnn = a->hnode_new; fa.Z3d = a->Z3d;
    fa.zbar3 = a->zbar3; fa.dt = a->dt; fa.remap = tune.remap(); fa.K = 0;
    fa.LO = LO; fa.gnode = (double2*)gnode; fa.adfh = adfh; fa.pf = pf; fa.adfv = adfv;
}

int launch_fused(const DevMesh& m, const NodeTab& tab, const int2* ekr, const fct_step_args* a,
                 int kw, const FusedTuning& tune, double* vflux, double* LO, double* gnode,
                 double* adfh, double* pf, double* adfv, hipStream_t s) {
    if (!tuning_ok(tune)) return set_error(FCT_E_INVALID, "bad launch-shape knobs");
    FusedArgs fa;
    base_args(fa, m, tab, ekr, a, tune, vflux, LO, gnode, adfh, pf, adfv);
    if (kw > FCT_KCHUNK) kw = FCT_KCHUNK;
    if (tab.dm == 6 && m.L <= 64) return launch_fused_t<1, 6>(fa, a, kw, tune, s);
    if (tab.dm == 6 && m.L <= 128) return launch_fused_t<2, 6>(fa, a, kw, tune, s);
    if (tab.dm == 8 && m.L <= 64) return launch_fused_t<1, 8>(fa, a, kw, tune, s);
    if (tab.dm == 8 && m.L <= 128) return launch_fused_t<2, 8>(fa, a, kw, tune, s);
    return set_error(FCT_E_UNSUPPORTED, "fused path supports node degree <= 8 and nl <= 128");
}

template <int NS>
static void launch_vflux_t(const FusedArgs& fa, const FusedTuning& t, hipStream_t s) {
    const dim3 blk(64 * t.wpb);
    const unsigned g = blocks_for(vflux_waves(fa.m.D, t), t.wpb);
    if (t.vflux_epw == 2) hipLaunchKernelGGL((kf_vflux<NS, 2>), dim3(g), blk, 0, s, fa);
    else hipLaunchKernelGGL((kf_vflux<NS, 4>), dim3(g), blk, 0, s, fa);
}

int launch_vflux_fused(const DevMesh& m, const NodeTab& tab, const int2* ekr, const double* uv,
                       const double* helem, const FusedTuning& tune, double* vflux,
                       hipStream_t s) {
    if (!tuning_ok(tune)) return set_error(FCT_E_INVALID, "bad launch-shape knobs");
    if (m.L > 128) return set_error(FCT_E_UNSUPPORTED, "fused edge flux supports nl <= 128");
    if (!m.D) return FCT_OK;
    FusedArgs fa{};
    fa.m = m; fa.tab = tab; fa.ekr = ekr; fa.uv = uv; fa.helem = helem; fa.vflux = vflux;
    fa.remap = tune.remap(); fa.K = 1;
    if (m.L <= 64) launch_vflux_t<1>(fa, tune, s);
    else launch_vflux_t<2>(fa, tune, s);
    return check_launch("kf_vflux");
}

template <int NS, int DM>
static int launch_upwind_t(FusedArgs& fa, const fct_step_args* a, const FusedTuning& t,
                           hipStream_t s) {
    const DevMesh& m = fa.m;
    fa.K = 1;
    if (m.D) launch_single<NS, DM>(fa, t, s, 0);
    int rc = check_launch("kf_vflux");
    if (rc) return rc;
    for (int j0 = 0; j0 < a->ntracers; j0 += FCT_KCHUNK) {
        const int K = (a->ntracers - j0) < FCT_KCHUNK ? (a->ntracers - j0) : FCT_KCHUNK;
        fa.K = K;
        for (int j = 0; j < FCT_KCHUNK; ++j) {
            const int jj = j0 + (j < K ? j : 0);
            fa.T[j] = a->T[jj]; fa.Told[j] = a->T_old[jj]; fa.Tnew[j] = a->T_new[jj];
        }
        const unsigned gn = blocks_for(node_waves(m.N, K), t.wpb);
        hipLaunchKernelGGL((kf_upwind<NS, DM>), dim3(gn), dim3(64 * t.wpb), 0, s, fa);
        if ((rc = check_launch("kf_upwind"))) return rc;
    }
    return FCT_OK;
}

int launch_upwind(const DevMesh& m, const NodeTab& tab, const int2* ekr, const fct_step_args* a,
                  const FusedTuning& tune, double* vflux, hipStream_t s) {
    if (!tuning_ok(tune)) return set_error(FCT_E_INVALID, "bad launch-shape knobs");
    FusedArgs fa;
    base_args(fa, m, tab, ekr, a, tune, vflux, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (tab.dm == 6 && m.L <= 64) return launch_upwind_t<1, 6>(fa, a, tune, s);
    if (tab.dm == 6 && m.L <= 128) return launch_upwind_t<2, 6>(fa, a, tune, s);
    if (tab.dm == 8 && m.L <= 64) return launch_upwind_t<1, 8>(fa, a, tune, s);
    if (tab.dm == 8 && m.L <= 128) return launch_upwind_t<2, 8>(fa, a, tune, s);
    return set_error(FCT_E_UNSUPPORTED, "upwind path supports node degree <= 8 and nl <= 128");
}

size_t fused_args_size() { return sizeof(FusedArgs); }

int fill_fused_args(void* out, const DevMesh& m, const NodeTab& tab, const int2* ekr,
                    const fct_step_args* a, const FusedTuning& tune, double* vflux, double* LO,
                    double* gnode, double* adfh, double* pf, double* adfv) {
    if (a->ntracers > FCT_KCHUNK)
        return set_error(FCT_E_UNSUPPORTED, "batched instances take at most %d tracers", FCT_KCHUNK);
    FusedArgs fa;
    base_args(fa, m, tab, ekr, a, tune, vflux, LO, gnode, adfh, pf, adfv);
    fa.K = a->ntracers;
    for (int j = 0; j < FCT_KCHUNK; ++j) {
        const int jj = j < fa.K ? j : 0;
        fa.T[j] = a->T[jj]; fa.Told[j] = a->T_old[jj]; fa.Tnew[j] = a->T_new[jj];
    }
    memcpy(out, &fa, sizeof(fa));
    return FCT_OK;
}

template <int NS, int DM>
static void launch_batch_t(const BatchGroup& g, const FusedTuning& t, hipStream_t s) {
    const dim3 blk(64 * t.wpb);
    if (g.nb_v) {
        if (t.vflux_epw == 2)
            hipLaunchKernelGGL((kb_vflux<NS, 2>), dim3(g.nb_v), blk, 0, s, g.desc, g.blk_v, t.remap());
        else
            hipLaunchKernelGGL((kb_vflux<NS, 4>), dim3(g.nb_v), blk, 0, s, g.desc, g.blk_v, t.remap());
    }
    const bool split = NS == 2 && g.blk_ns;
    if (g.nb_pa) hipLaunchKernelGGL((kbp_pass1<DM, 2>), dim3(g.nb_pa), blk, 0, s, g.desc, g.blk_pa, t.remap());
    if (g.nb_p1) hipLaunchKernelGGL((kbp_pass1<DM, 1>), dim3(g.nb_p1), blk, 0, s, g.desc, g.blk_p1, t.remap());
    if (split) {
        if (g.nb_ns) hipLaunchKernelGGL((kbl_pass1<1, DM>), dim3(g.nb_ns), blk, 0, s, g.desc, g.blk_ns, t.remap(), 0);
        if (g.nb_nd) hipLaunchKernelGGL((kbl_pass1<2, DM>), dim3(g.nb_nd), blk, 0, s, g.desc, g.blk_nd, t.remap(), 1);
    } else if (g.nb_n) {
        hipLaunchKernelGGL((kb_pass1<NS, DM>), dim3(g.nb_n), blk, 0, s, g.desc, g.blk_n, t.remap());
    }
    if (g.nb_e2 || g.nb_e1) {
#define FCT_KBE(KC, TPE)                                                                            \
    do {                                                                                            \
        if (g.nb_e2) hipLaunchKernelGGL((kbe_adfh<2, KC, TPE>), dim3(g.nb_e2), dim3(FCT_EP * TPE), 0, s, \
                                        g.desc, g.blk_e2, t.remap());                               \
        if (g.nb_e1) hipLaunchKernelGGL((kbe_adfh<1, KC, TPE>), dim3(g.nb_e1), dim3(FCT_EP * TPE), 0, s, \
                                        g.desc, g.blk_e1, t.remap());                               \
    } while (0)
        if (t.adfh_kc == 4) {
            if (t.adfh_tpe == 1) FCT_KBE(4, 1);
            else if (t.adfh_tpe == 2) FCT_KBE(4, 2);
            else FCT_KBE(4, 4);
        } else {
            if (t.adfh_tpe == 1) FCT_KBE(8, 1);
            else if (t.adfh_tpe == 2) FCT_KBE(8, 2);
            else FCT_KBE(8, 4);
        }
#undef FCT_KBE
    }
    if (g.nb_a) {
        if (t.adfh_epw == 1)
            hipLaunchKernelGGL((kb_adfh<NS, 1>), dim3(g.nb_a), blk, 0, s, g.desc, g.blk_a, t.remap());
        else
            hipLaunchKernelGGL((kb_adfh<NS, 2>), dim3(g.nb_a), blk, 0, s, g.desc, g.blk_a, t.remap());
    }
    if (g.nb_pa) hipLaunchKernelGGL((kbp_pass2<DM, 2>), dim3(g.nb_pa), blk, 0, s, g.desc, g.blk_pa, t.remap());
    if (g.nb_p1) hipLaunchKernelGGL((kbp_pass2<DM, 1>), dim3(g.nb_p1), blk, 0, s, g.desc, g.blk_p1, t.remap());
    if (split) {
        if (g.nb_ns) hipLaunchKernelGGL((kbl_pass2<1, DM>), dim3(g.nb_ns), blk, 0, s, g.desc, g.blk_ns, t.remap(), 0);
        if (g.nb_nd) hipLaunchKernelGGL((kbl_pass2<2, DM>), dim3(g.nb_nd), blk, 0, s, g.desc, g.blk_nd, t.remap(), 1);
    } else if (g.nb_n) {
        hipLaunchKernelGGL((kb_pass2<NS, DM>), dim3(g.nb_n), blk, 0, s, g.desc, g.blk_n, t.remap());
    }
    if (g.nb_pa) hipLaunchKernelGGL((kbp_pass3<DM, 2>), dim3(g.nb_pa), blk, 0, s, g.desc, g.blk_pa, t.remap());
    if (g.nb_p1) hipLaunchKernelGGL((kbp_pass3<DM, 1>), dim3(g.nb_p1), blk, 0, s, g.desc, g.blk_p1, t.remap());
    if (split) {
        if (g.nb_ns) hipLaunchKernelGGL((kbl_pass3<1, DM>), dim3(g.nb_ns), blk, 0, s, g.desc, g.blk_ns, t.remap(), 0);
        if (g.nb_nd) hipLaunchKernelGGL((kbl_pass3<2, DM>), dim3(g.nb_nd), blk, 0, s, g.desc, g.blk_nd, t.remap(), 1);
    } else if (g.nb_n) {
        hipLaunchKernelGGL((kb_pass3<NS, DM>), dim3(g.nb_n), blk, 0, s, g.desc, g.blk_n, t.remap());
    }
}

int launch_batch_group(const BatchGroup& g, const FusedTuning& t, hipStream_t s) {
    if (!tuning_ok(t)) return set_error(FCT_E_INVALID, "bad launch-shape knobs");
    if (g.dm == 6 && g.ns == 1) launch_batch_t<1, 6>(g, t, s);
    else if (g.dm == 6 && g.ns == 2) launch_batch_t<2, 6>(g, t, s);
    else if (g.dm == 8 && g.ns == 1) launch_batch_t<1, 8>(g, t, s);
    else if (g.dm == 8 && g.ns == 2) launch_batch_t<2, 8>(g, t, s);
    else return set_error(FCT_E_UNSUPPORTED, "batch group (%d slots, degree %d)", g.ns, g.dm);
    return check_launch("batched fused passes");
}

}  // namespace fct
