// Shared device-side definitions of the FCT library: the device mesh view and the
// small helpers every kernel uses. Constants are those of
// fesom_jax/tracer_adv.py:44-46, 192-193 and fesom_jax/config.py:41.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "fesom_fct.h"

namespace fct {

constexpr double R_EARTH = 6367500.0;
constexpr double AB_OLD = -(0.5 + 0.1);
constexpr double AB_NEW = 1.5 + 0.1;
constexpr double FCT_FLUX_EPS = 1e-16;
constexpr double FCT_BIG = 1e3;
constexpr int FCT_KCHUNK = 8;   // tracers per fused launch (kernel-argument pointer slots)

// Device-resident view of the mesh, passed to every kernel by value.
struct DevMesh {
    int N, E, D, L;
    const int2* edges;        // [D] (n1, n2)
    const int2* edge_tri;     // [D] (el1, el2), -1 = none
    const int2* updn;         // [D] (up, dn) triangles, clamped to >= 0
    const int2* utri_rng;     // [D] level range [lo, hi) where the up/dn gradients apply
    const double4* cross;     // [D] (dx1, dy1, dx2, dy2) metres
    const double2* edge_c;    // [D] (edx * a, edy * R_EARTH), a = R cos averaged
    const int* en;            // [E*3]
    const double* gs;         // [E*6]
    const double* earea;      // [E]
    const int2* elev;         // [E] layer range [ulevels-1, nlevels-1)
    const int2* nlev;         // [N] (nzmin, nzmax) = (ulevels_nod2D-1, nlevels_nod2D-1)
    const int* ecsr_off;      // [N+1]
    const int* ecsr;          // 2*edge + side, increasing per node
    const int* ncsr_off;      // [N+1]
    const int* ncsr;          // element ids, increasing per node
    const double* area;       // [N*L]
    const double* areasvol;   // [N*L]
    const double* Zs;         // [4*L] static QR4C stencil depths Z[k], Z[k-1], Z[k-2], Z[k+1]
    const double* zbar;       // [L]
    // fused path: (area, areasvol) at every node level and (Z[k], zbar[k]) of the static
    // stencil, for the passes that need both of a pair: one 16-byte load instead of two
    // 8-byte ones
    const double2* aav;       // [N*L]
    const double2* zzb;       // [L]
    // fused-path edge tables (nl <= 128): the layer ranges of an edge's two triangles as
    // bytes lo1|hi1|lo2|hi2 (a missing triangle: empty), and the up and down gradient
    // triangles: the two vertices other than the edge's first node (up) and second node
    // (down), in vertex order, (u_a, u_b, d_a, d_b), then (cu | cd << 2, -, -, -) with
    // cu, cd the vertex slots of the edge's own nodes
    const int* etr;           // [D]
    const int4* udv;          // [2*D]
};

// Edge patches of the fused path. The plan groups the nodes into compact patches grown
// breadth-first over the mesh graph (whatever the caller's numbering), and each patch
// takes the edges whose first node it holds: at most FCT_EP edges and FCT_PU distinct
// nodes (the edges' ends and their up/down triangle vertices) per patch. A patch
// kernel loads the columns of its node list once per level chunk into LDS and runs
// one edge per thread over the chunk's levels.
constexpr int FCT_EP = 256;   // edges per patch (threads per workgroup)
constexpr int FCT_PU = 168;   // node slots per patch
struct EdgePatches {
    int np = 0;                   // patches
    const int4* ptab = nullptr;   // [np] (first task, node-list offset,
                                  //  edges | nodes << 16, klo | khi << 8 | kg << 16):
                                  //  [klo, khi) = the union of the patch edges' wet level
                                  //  ranges, kg = the first level at which an edge reads
                                  //  node-averaged gradients (khi if none does)
    const int* pnodes = nullptr;  // node list of each patch
    const int* elist = nullptr;   // [D] edge of each task, patch by patch
    const int2* eslot = nullptr;  // [D] per task: LDS slots of (n1, n2, u_a, u_b) as bytes;
                                  //  (d_a, d_b) as bytes | cu << 16 | cd << 18
};

// Per-node stencil table of the fused path (dm slots per node).
struct NodeTab {
    int dm = 0;               // slots per node (6 or 8); 0 = mesh not supported
    const int4* hdr;          // [N] (nzmin, nzmax, incident edges | glo << 8 | ghi << 16,
                              //  incident elements); [glo, ghi) = levels at which the
                              //  adf_h pass reads the node's gnode (empty = none)
    const int4* edge;         // [N*dm] (2*edge + side, other node, layer ranges of the
                              //  edge's triangles as bytes lo1|hi1|lo2|hi2, layer range of
                              //  element slot q as bytes lo|hi), edges in reference order
    const int4* elem;         // [N*dm] (element, v0, v1, v2), increasing element id
    // nl > 64 only: the nodes whose wet layers end above level 64 (`nshallow` of them, in
    // node order), then the others; passes 2 and 3 take each set in a launch of its own,
    // the first with one 64-level slot per wavefront and fewer registers
    const int* nlist = nullptr;
    int nshallow = 0;
    // slots past a node's counts repeat its first slot, so a kernel can load every slot
    // without waiting for the counts
    EdgePatches ep;           // edge patches of the adf_h patch kernel
    // order in which the vertical-flux kernel visits the edges (NULL = id order); the
    // patches' edge order when the caller's numbering is not local
    const int* vorder = nullptr;
};

// Host side of the edge patches (fct_locality.hip), tables as in EdgePatches, and the
// breadth-first node order of the patches' balls. `udv` and
// `has_updn` are the up/down triangle vertex table and whether the edge has both
// triangles, `utri` the level range in which an edge uses its up/down triangles'
// gradients.
struct LocalityHost {
    std::vector<int> node_order;    // nodes ball by ball (breadth-first), balls in order
    std::vector<int4> ptab;
    std::vector<int> pnodes, elist;
    std::vector<int2> eslot;
};
int build_locality(int N, int D, const int2* edges, const int* eoff, const int* ecsr,
                   const int4* udv, const unsigned char* has_updn, const int2* ekr,
                   const int2* utri, LocalityHost& out);

__device__ __forceinline__ bool in_layer(int2 r, int k) { return k >= r.x && k < r.y; }
__device__ __forceinline__ bool in_iface(int2 r, int k) { return k >= r.x && k <= r.y; }
__device__ __forceinline__ double nz_safe(double x) { return x == 0.0 ? 1.0 : x; }

// One row of adv_tra_vert_impl's tridiagonal system (fesom_jax/tracer_adv.py:487-529)
// at level k of the column starting at b0: sub/super-diagonal a, c, diagonal b, the
// advective part b_adv of the diagonal, and the pieces the reverse mode needs.
struct VimplRow {
    double a, b, c, b_adv, v_top, v_bot, wu, wd;
    bool surf, bot, valid;
};
__device__ __forceinline__ VimplRow vimpl_row(const DevMesh& m, const double* __restrict__ w_i,
                                              const double* __restrict__ hnn, double dt,
                                              size_t b0, int k, int2 nr) {
    const int L = m.L;
    const size_t i = b0 + k;
    VimplRow r;
    const double av = m.areasvol[i];
    const double safe_av = av > 0.0 ? av : 1.0;
    r.v_top = dt * m.area[i] / safe_av;
    r.v_bot = dt * (k + 1 < L ? m.area[i + 1] : 0.0) / safe_av;
    r.wu = w_i[i];
    r.wd = k + 1 < L ? w_i[i + 1] : 0.0;
    r.surf = k == nr.x;
    r.bot = k == nr.y - 1;
    r.valid = in_layer(nr, k);
    const double a_adv = (r.surf ? 0.0 : fmin(0.0, r.wu)) * r.v_top;
    const double b_top = (r.surf ? r.wu : fmax(0.0, r.wu)) * r.v_top;
    const double b_bot = (r.bot ? 0.0 : -fmin(0.0, r.wd)) * r.v_bot;
    const double c_adv = (r.bot ? 0.0 : -fmax(0.0, r.wd)) * r.v_bot;
    r.b_adv = b_top + b_bot;
    r.a = r.valid ? a_adv : 0.0;
    r.c = r.valid ? c_adv : 0.0;
    r.b = r.valid ? hnn[i] + r.b_adv : 1.0;
    return r;
}

// Host helpers (defined in fct_api.hip).
int set_error(int code, const char* fmt, ...);
int check_launch(const char* what);

// Phased launchers (fct_phased.hip).
int launch_vflux(const DevMesh& m, const fct_step_args* a, double* vflux, hipStream_t s);
int launch_lo(const DevMesh& m, const fct_step_args* a, int j, const double* vflux,
              double* LO, double* scr0, double* scr1, hipStream_t s);
int launch_grad(const DevMesh& m, const fct_step_args* a, int j, double* tr, hipStream_t s);
// adv_tra_vert_impl of X in place (cp, dp: [N*L] scratch)
int launch_vert_impl(const DevMesh& m, const double* w_i, const double* hnn, double dt,
                     double* X, double* cp, double* dp, hipStream_t s);
int launch_pm(const DevMesh& m, const fct_step_args* a, int j, const double* vflux,
              const double* LO, const double* tr, double* gnode, double* adfh, double* adfv,
              double* tvmax, double* tvmin, double* fplus, double* fminus, hipStream_t s);
int launch_final(const DevMesh& m, const fct_step_args* a, int j, const double* LO,
                 const double* adfh, const double* adfv, const double* tvmax,
                 const double* tvmin, const double* fplus, const double* fminus, double* pf,
                 double* mf, hipStream_t s);

// Scratch of the single-stage entry point (plan workspace).
struct StageScratch {
    double *vflux = nullptr, *gnode = nullptr, *tvmax = nullptr, *tvmin = nullptr;
    double *fplus = nullptr, *fminus = nullptr, *pf = nullptr, *mf = nullptr;
};
int launch_stage(const DevMesh& m, const StageScratch& w, int stage, const double* const* in,
                 double* const* out, double dt, hipStream_t s);
int launch_tracer_ab(const double* T, const double* Told, double* out, int64_t count,
                     hipStream_t s);

int launch_fac(const DevMesh& m, const fct_step_args* a, const double* LO, const double* tvmax,
               const double* tvmin, const double* fplus, const double* fminus, double* pf,
               double* mf, hipStream_t s);

// Workspace of the adjoint (fct_adjoint.hip): the forward tape of one tracer (dense
// phased intermediates), cotangent scratch, and the element-side CSR lists.
struct AdjointWork {
    // tape
    double *vflux, *LO, *trxy, *gnode, *adfh, *adfv, *tvmax, *tvmin, *fplus, *fminus, *pf, *mf;
    // node scratch [N*L] (g_num is [N*L*2])
    double *gLO, *gx, *ghsc, *g_pf, *g_mf, *g_fp, *g_fm, *g_fmax, *g_fmin, *g_adfv, *gA;
    double *g_fv, *g_tvx, *g_tvn, *sh_x, *sh_n, *g_div, *g_num, *gw;
    // edge scratch: g_adfh [D*L], eA and eT [D*L*2], eud [D*L*4], g_vf [D*L]
    double *g_adfh, *eA, *eT, *eud, *g_vf;
    // element scratch: g_tr [E*L*2]
    double* g_tr;
    // element -> (edge, up/down role) and element -> (edge, side) lists
    const int *ud_off, *ud_ent, *et_off, *et_ent;
    // implicit split (allocated on the first backward with w_i): LO before the
    // implicit solve, and the w_i cotangent when the caller does not take it
    double *LOx = nullptr, *gwi = nullptr;
};

int launch_backward(const DevMesh& m, const AdjointWork& w, const fct_step_args* a,
                    const double* const* g_out, double* const* gT, double* const* gTold,
                    double* g_uv, double* g_w, double* g_wi, hipStream_t s);

// Launch-shape knobs of the fused path (fct_plan_set_tuning / fct_batch_set_tuning);
// results do not depend on them.
struct FusedTuning {
    int xcd = 8;        // remap blocks so each of `xcd` XCDs sweeps its own tile ranges
                        // (1 = hardware order)
    int xchunk = 8;     // tiles per contiguous range of one XCD (0 = one range per XCD)
    int wpb = 4;        // wavefronts per workgroup (1, 2 or 4)
    int vflux_epw = 4;  // edges per wavefront in kf_vflux (2 or 4)
    int adfh_epw = 2;   // edges per wavefront in kf_adfh (1 or 2)
    int adfh_kc = 8;    // adf_h pass: levels per chunk of the edge-patch kernel (4 or 8),
                        // 0 = one wavefront per edge column (kf_adfh)
    int adfh_tpe = 1;   // edge-patch kernel: threads per edge (1, 2 or 4), which split
                        // the chunk's levels
    int w_vflux = 1;    // compute_w from the edge volume flux (1) or per node (0)
    int pair = 1;       // two tracers: shallow node passes take both in one wavefront
    // the remap as the kernels take it: XCD count in the low 8 bits, chunk above
    int remap() const { return xcd | (xchunk << 8); }
};

struct FusedArgs;

// Batched launch of several plans' fused steps: for every pass kind, one launch over
// the concatenated task ranges of the instances of one (slots, degree) class.
struct BatchGroup {
    int ns = 1, dm = 6;
    int count = 0;                 // instances in the group
    const FusedArgs* desc = nullptr;   // [count] device descriptors
    const int2* blk_v = nullptr;   // per workgroup: (instance, first wave) of the vflux pass
    const int2* blk_n = nullptr;   // ... of the node passes
    const int2* blk_a = nullptr;   // ... of the adfh pass
    // nl > 64: node passes split by column depth, waves over each instance's shallow
    // and deep node lists (NodeTab::nlist)
    const int2* blk_ns = nullptr;
    const int2* blk_nd = nullptr;
    // several tracers per wavefront (instances with K >= 2 on 6-slot meshes): waves over
    // each such instance's shallow node list, per pair of tracer slots (blk_pa) and for
    // the last slot of an odd K (blk_p1), the instance field carrying the first slot in
    // its high 16 bits; the deep nodes (nl > 64) stay in blk_nd
    const int2* blk_pa = nullptr;
    const int2* blk_p1 = nullptr;
    // adf_h by edge patches: per workgroup (instance, patch | first tracer << 24), two
    // tracers per thread (e2) or one (e1); instances without patches stay in blk_a
    const int2* blk_e2 = nullptr;
    const int2* blk_e1 = nullptr;
    unsigned nb_v = 0, nb_n = 0, nb_a = 0, nb_ns = 0, nb_nd = 0, nb_pa = 0, nb_p1 = 0, nb_e2 = 0,
             nb_e1 = 0;
};

// Fused launcher (fct_fused.hip): kw = tracers per chunk (workspace slots).
int launch_fused(const DevMesh& m, const NodeTab& tab, const int2* ekr, const fct_step_args* a,
                 int kw, const FusedTuning& tune, double* vflux, double* LO, double* gnode,
                 double* adfh, double* pf, double* adfv, hipStream_t s);

// Upwind-only step (fct_fused.hip).
int launch_upwind(const DevMesh& m, const NodeTab& tab, const int2* ekr, const fct_step_args* a,
                  const FusedTuning& tune, double* vflux, hipStream_t s);

// Host-side descriptor of one instance's fused step (K <= kw tracers, one chunk).
size_t fused_args_size();
int fill_fused_args(void* out, const DevMesh& m, const NodeTab& tab, const int2* ekr,
                    const fct_step_args* a, const FusedTuning& tune, double* vflux,
                    double* LO, double* gnode, double* adfh, double* pf, double* adfv);
int launch_batch_group(const BatchGroup& g, const FusedTuning& tune, hipStream_t s);
// tasks (wavefronts) of instance with D edges, N nodes, K tracers for each pass kind
inline int64_t vflux_waves(int D, const FusedTuning& t) { return (D + t.vflux_epw - 1) / t.vflux_epw; }
inline int64_t node_waves(int N, int K) { return (int64_t)N * K; }
inline int64_t adfh_waves(int D, int K, const FusedTuning& t) {
    return (int64_t)((D + t.adfh_epw - 1) / t.adfh_epw) * K;
}

// Sea-ice FCT of the three ice tracers (fct_ice.hip).
struct IceCall {
    double dt = 0.0, diff = 0.0, gamma = 0.0;
    const double* u = nullptr;
    const double* v = nullptr;
    const double* X[3] = {nullptr, nullptr, nullptr};   // a_ice, m_ice, m_snow
    double* out[3] = {nullptr, nullptr, nullptr};
};
size_t ice_workspace_doubles(int N);
size_t ice_table_doubles(int E);
// the per-element table of the sea-ice passes (vertices, open flag, gradient
// coefficients, area, sq[e] = sqrt(elem_area[e] / 2e8) computed on the host, correctly
// rounded), built once per plan into `tab`
int launch_ice_tables(const DevMesh& m, const double* sq, double* tab, hipStream_t s);
int launch_ice_fct(const DevMesh& m, const IceCall& c, double* work, const double* tab,
                   hipStream_t s);

// Vertical velocity from continuity (fct_ale.hip).
// the FCT's edge volume flux alone (kf_vflux) into vflux [D][L]; fused node tables only
int launch_vflux_fused(const DevMesh& m, const NodeTab& tab, const int2* ekr, const double* uv,
                       const double* helem, const FusedTuning& t, double* vflux, hipStream_t s);
// compute_w's divergence from that edge flux (exactly minus the per-edge term)
int launch_compute_w_vf(const DevMesh& m, const NodeTab& tab, const int2* ekr,
                        const double* vflux, double* w, hipStream_t s);
int launch_compute_w(const DevMesh& m, const double* uv, const double* helem, double* w,
                     hipStream_t s);

// Implicit vertical diffusion of K <= 2 tracers sharing one matrix (fct_vdiff.hip).
struct VdiffCall {
    int K = 0;
    double dt = 0.0;
    const double* Kv = nullptr;
    const double* hnn = nullptr;
    const double* Z3d = nullptr;
    const double* T[2] = {nullptr, nullptr};
    const double* bc[2] = {nullptr, nullptr};
    const double* sw[2] = {nullptr, nullptr};
    double* Tn[2] = {nullptr, nullptr};
    int floor_j = -1;               // index within the call, -1 = none
    double floor_v = 0.0;
};
// Workgroups the GPU holds at once over every kernel variant (cus compute units).
int vdiff_max_grid(int cus, int* out);
size_t vdiff_scratch_doubles(int L, int grid);
int launch_vdiff(const DevMesh& m, const VdiffCall& c, int max_grid, double* scratch,
                 hipStream_t s);
// Device address of the fixed scratch of fct_impl_vert_diff_static on `device` (the
// current device), looked up once per device.
int vdiff_static_scratch(int device, double** out);

}  // namespace fct
