// C ABI of libfesom_fct.so: plan construction (topology, CSR lists in the
// reference's scatter order, geometry, workspace) and the forward entry points.
// The declarations and their reference counterparts are in include/fesom_fct.h.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>

#include "fct_common.h"

namespace fct {

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int check_launch(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return set_error(FCT_E_HIP, "%s: launch failed: %s", what, hipGetErrorString(e));
    return FCT_OK;
}

}  // namespace fct

using namespace fct;

#define HIPCHK(expr)                                                                  \
    do {                                                                              \
        const hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess)                                                         \
            return set_error(FCT_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// Device tables of one mesh (topology, geometry, stencil tables, edge patches): built
// by fct_plan_create, shared by the plans fct_plan_clone makes from it, freed with the
// last of them.
struct TableSet {
    int refs = 1;                 // guarded by table_mutex
    std::vector<void*> allocs;
    int64_t bytes = 0;
};
static std::mutex table_mutex;

struct fct_plan {
    int device = -1;
    DevMesh m{};
    TableSet* tables = nullptr;
    std::vector<void*> allocs;    // the plan's own workspace
    int64_t bytes = 0;
    // workspace
    double *vflux = nullptr, *LO = nullptr, *trxy = nullptr, *gnode = nullptr;
    double *adfh = nullptr, *adfv = nullptr, *tvmax = nullptr, *tvmin = nullptr;
    double *fplus = nullptr, *fminus = nullptr, *pf = nullptr, *mf = nullptr;
    // fused path: per-edge wet level range and per-tracer workspace (kw tracers)
    const int2* ekr = nullptr;
    int kw = 0;
    FusedTuning tune;
    NodeTab tab;
    // fpf holds the limiter factors as (plus, minus) pairs; fadfv the vertical
    // antidiffusive flux pass 2 computes and pass 3 reads
    double *fLO = nullptr, *fgnode = nullptr, *fadfh = nullptr, *fpf = nullptr, *fadfv = nullptr;
    // adjoint workspace (allocated on the first fct_backward)
    bool adj_ready = false;
    AdjointWork aw{};
    std::vector<int> host_edge_tri, host_updn;   // [2D] each, -1 sentinels (adjoint lists)
    // implicit vertical diffusion scratch (allocated on the first call)
    double* vd_scratch = nullptr;
    int vd_grid = 0;
    // sea-ice FCT workspace and element table (allocated on the first call)
    double* ice_work = nullptr;
    double* ice_tab = nullptr;
};

namespace {

template <class T>
int dev_alloc(fct_plan* p, T** out, size_t count) {
    *out = nullptr;
    if (count == 0) count = 1;
    void* ptr = nullptr;
    const hipError_t e = hipMalloc(&ptr, count * sizeof(T));
    if (e != hipSuccess)
        return set_error(FCT_E_NOMEM, "hipMalloc(%zu bytes) failed: %s", count * sizeof(T),
                         hipGetErrorString(e));
    p->allocs.push_back(ptr);
    p->bytes += (int64_t)(count * sizeof(T));
    *out = static_cast<T*>(ptr);
    return FCT_OK;
}

template <class T>
int dev_upload(fct_plan* p, const T** out, const T* host, size_t count) {
    T* d = nullptr;
    int rc = dev_alloc(p, &d, count);
    if (rc) return rc;
    if (count) HIPCHK(hipMemcpy(d, host, count * sizeof(T), hipMemcpyHostToDevice));
    *out = d;
    return FCT_OK;
}

void free_plan(fct_plan* p) {
    if (!p) return;
    int cur = -1;
    if (p->device >= 0 && hipGetDevice(&cur) == hipSuccess && cur != p->device)
        (void)hipSetDevice(p->device);
    for (void* ptr : p->allocs) (void)hipFree(ptr);
    if (p->tables) {
        bool last;
        {
            std::lock_guard<std::mutex> g(table_mutex);
            last = --p->tables->refs == 0;
        }
        if (last) {
            for (void* ptr : p->tables->allocs) (void)hipFree(ptr);
            delete p->tables;
        }
    }
    if (cur >= 0 && cur != p->device) (void)hipSetDevice(cur);
    delete p;
}

// The fused path's workspace of a plan (kw tracer slots); the edge passes write only
// each edge's wet levels, so vflux and adf_h start at zero, the value the node passes
// must read elsewhere.
int alloc_fused_workspace(fct_plan* p) {
    const size_t NL = (size_t)p->m.N * p->m.L, DL = (size_t)p->m.D * p->m.L;
    int rc;
    if ((rc = dev_alloc(p, &p->vflux, DL)) || (rc = dev_alloc(p, &p->fLO, NL * p->kw)) ||
        (rc = dev_alloc(p, &p->fgnode, 2 * NL * p->kw)) ||
        (rc = dev_alloc(p, &p->fadfh, DL * p->kw)) ||
        (rc = dev_alloc(p, &p->fpf, 2 * NL * p->kw)) ||
        (rc = dev_alloc(p, &p->fadfv, NL * p->kw)))
        return rc;
    if (hipMemset(p->vflux, 0, DL * sizeof(double)) != hipSuccess ||
        hipMemset(p->fadfh, 0, DL * p->kw * sizeof(double)) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess)
        return set_error(FCT_E_HIP, "workspace initialisation failed");
    return FCT_OK;
}

int validate_mesh(const fct_mesh_desc* d) {
    if (!d) return set_error(FCT_E_INVALID, "mesh descriptor is NULL");
    if (d->nod2D <= 0 || d->elem2D <= 0 || d->edge2D <= 0 || d->nl < 2)
        return set_error(FCT_E_INVALID, "bad mesh sizes nod2D=%d elem2D=%d edge2D=%d nl=%d",
                         d->nod2D, d->elem2D, d->edge2D, d->nl);
    const int64_t big = (int64_t)d->edge2D * d->nl * 2;
    if ((int64_t)d->elem2D * d->nl * 2 > big || big > (int64_t)INT32_MAX * 2)
        return set_error(FCT_E_UNSUPPORTED, "mesh too large for this build");
    const void* ptrs[] = {d->edges, d->edge_tri, d->edge_cross_dxdy, d->edge_dxdy,
                          d->edge_up_dn_tri, d->elem_nodes, d->gradient_sca, d->elem_area,
                          d->elem_cos, d->nlevels, d->ulevels, d->nlevels_nod2D,
                          d->nlevels_nod2D_min, d->ulevels_nod2D, d->ulevels_nod2D_max,
                          d->area, d->areasvol, d->Z, d->zbar};
    for (const void* q : ptrs)
        if (!q) return set_error(FCT_E_INVALID, "mesh descriptor has a NULL array");
    const int N = d->nod2D, E = d->elem2D, D = d->edge2D, L = d->nl;
    for (int e = 0; e < D; ++e) {
        const int a = d->edges[2 * e], b = d->edges[2 * e + 1];
        if (a < 0 || a >= N || b < 0 || b >= N || a == b)
            return set_error(FCT_E_INVALID, "edge %d has bad nodes (%d, %d)", e, a, b);
        const int t1 = d->edge_tri[2 * e], t2 = d->edge_tri[2 * e + 1];
        if (t1 < -1 || t1 >= E || t2 < -1 || t2 >= E)
            return set_error(FCT_E_INVALID, "edge %d has bad edge_tri (%d, %d)", e, t1, t2);
        const int u = d->edge_up_dn_tri[2 * e], w = d->edge_up_dn_tri[2 * e + 1];
        if (u < -1 || u >= E || w < -1 || w >= E)
            return set_error(FCT_E_INVALID, "edge %d has bad edge_up_dn_tri (%d, %d)", e, u, w);
    }
    for (int e = 0; e < E; ++e) {
        for (int j = 0; j < 3; ++j) {
            const int v = d->elem_nodes[3 * e + j];
            if (v < 0 || v >= N)
                return set_error(FCT_E_INVALID, "element %d has bad node %d", e, v);
        }
        // every level bound in [1, nl] and top <= bottom (the fused tables pack them into
        // bytes, so an out-of-range value must not get that far)
        if (d->ulevels[e] < 1 || d->ulevels[e] > L || d->nlevels[e] < d->ulevels[e] ||
            d->nlevels[e] > L)
            return set_error(FCT_E_INVALID, "element %d has bad levels (%d, %d), nl = %d", e,
                             d->ulevels[e], d->nlevels[e], L);
    }
    for (int n = 0; n < N; ++n) {
        const int lv[4] = {d->ulevels_nod2D[n], d->ulevels_nod2D_max[n], d->nlevels_nod2D_min[n],
                           d->nlevels_nod2D[n]};
        for (int q = 0; q < 4; ++q)
            if (lv[q] < 1 || lv[q] > L)
                return set_error(FCT_E_INVALID, "node %d has a level bound %d outside [1, %d]", n,
                                 lv[q], L);
        if (d->nlevels_nod2D[n] < d->ulevels_nod2D[n] ||
            d->ulevels_nod2D_max[n] < d->ulevels_nod2D[n] ||
            d->nlevels_nod2D_min[n] > d->nlevels_nod2D[n])
            return set_error(FCT_E_INVALID, "node %d has inconsistent levels (%d, %d, %d, %d)", n,
                             lv[0], lv[1], lv[2], lv[3]);
    }
    if (d->nod_in_elem2D_offsets) {
        const int32_t* off = d->nod_in_elem2D_offsets;
        if (off[0] != 0)
            return set_error(FCT_E_INVALID, "nod_in_elem2D_offsets[0] = %d, expected 0", off[0]);
        for (int n = 0; n < N; ++n)
            if (off[n + 1] < off[n])
                return set_error(FCT_E_INVALID, "nod_in_elem2D_offsets decrease at node %d", n);
    }
    return FCT_OK;
}

// Mesh properties the fused node passes rely on (checked once; the phased path is
// used when they fail):
//  - every triangle adjacent to an edge (n, p) contains n and p, and every triangle
//    around n is adjacent to the two edges of n it contains, so the vertices of the
//    triangles around n are n and the far ends of n's edges;
//  - a triangle's layer range lies inside each vertex's layer range, so a vertex of a
//    triangle that is wet at level k is itself wet at k.
bool fused_mesh_ok(const fct_mesh_desc* d, const std::vector<int>& eoff,
                   const std::vector<int>& ecsr, const std::vector<int>& noff,
                   const std::vector<int>& ncsr) {
    const int N = d->nod2D;
    auto has = [&](int el, int v) {
        return d->elem_nodes[3 * el] == v || d->elem_nodes[3 * el + 1] == v ||
               d->elem_nodes[3 * el + 2] == v;
    };
    for (int n = 0; n < N; ++n) {
        for (int q = eoff[n]; q < eoff[n + 1]; ++q) {
            const int e = ecsr[q] >> 1;
            const int p = (ecsr[q] & 1) ? d->edges[2 * e] : d->edges[2 * e + 1];
            for (int side = 0; side < 2; ++side) {
                const int el = d->edge_tri[2 * e + side];
                if (el >= 0 && !(has(el, n) && has(el, p))) return false;
            }
        }
        for (int q = noff[n]; q < noff[n + 1]; ++q) {
            const int el = ncsr[q];
            for (int j = 0; j < 3; ++j) {
                const int v = d->elem_nodes[3 * el + j];
                if (v == n) continue;
                bool found = false;
                for (int r = eoff[n]; r < eoff[n + 1] && !found; ++r) {
                    const int e = ecsr[r] >> 1;
                    const int p = (ecsr[r] & 1) ? d->edges[2 * e] : d->edges[2 * e + 1];
                    found = p == v && (d->edge_tri[2 * e] == el || d->edge_tri[2 * e + 1] == el);
                }
                if (!found) return false;
            }
            if (d->nlevels[el] > d->ulevels[el] &&
                (d->ulevels_nod2D[n] > d->ulevels[el] || d->nlevels_nod2D[n] < d->nlevels[el]))
                return false;
        }
    }
    return true;
}

// Workspace of the phased path, allocated on the first call that takes it (the fused
// path, which batches and most callers use, never needs it).
int ensure_phased(fct_plan* p) {
    if (p->LO) return FCT_OK;
    const size_t NL = (size_t)p->m.N * p->m.L, DL = (size_t)p->m.D * p->m.L,
                 EL = (size_t)p->m.E * p->m.L;
    int rc;
    if ((rc = dev_alloc(p, &p->trxy, 2 * EL)) || (rc = dev_alloc(p, &p->gnode, 2 * NL)) ||
        (rc = dev_alloc(p, &p->adfh, DL)) || (rc = dev_alloc(p, &p->adfv, NL)) ||
        (rc = dev_alloc(p, &p->tvmax, NL)) || (rc = dev_alloc(p, &p->tvmin, NL)) ||
        (rc = dev_alloc(p, &p->fplus, NL)) || (rc = dev_alloc(p, &p->fminus, NL)) ||
        (rc = dev_alloc(p, &p->pf, NL)) || (rc = dev_alloc(p, &p->mf, NL)) ||
        (rc = dev_alloc(p, &p->LO, NL)))
        return rc;
    return FCT_OK;
}

// Adjoint workspace: the phased tape plus cotangent scratch and the element-side lists
// (element -> (edge, up/down role) for the interior edges' gradient triangles, element
// -> (edge, side) for the edges' two triangles), in increasing edge order.
int ensure_adjoint(fct_plan* p) {
    if (p->adj_ready) return FCT_OK;
    int rc = ensure_phased(p);
    if (rc) return rc;
    const DevMesh& m = p->m;
    const size_t NL = (size_t)m.N * m.L, DL = (size_t)m.D * m.L, EL = (size_t)m.E * m.L;
    AdjointWork& w = p->aw;
    w.vflux = p->vflux; w.LO = p->LO; w.trxy = p->trxy; w.gnode = p->gnode; w.adfh = p->adfh;
    w.adfv = p->adfv; w.tvmax = p->tvmax; w.tvmin = p->tvmin; w.fplus = p->fplus;
    w.fminus = p->fminus; w.pf = p->pf; w.mf = p->mf;
    double** node[] = {&w.gLO, &w.gx, &w.ghsc, &w.g_pf, &w.g_mf, &w.g_fp, &w.g_fm, &w.g_fmax,
                       &w.g_fmin, &w.g_adfv, &w.gA, &w.g_fv, &w.g_tvx, &w.g_tvn, &w.sh_x,
                       &w.sh_n, &w.g_div, &w.gw};
    for (double** q : node)
        if ((rc = dev_alloc(p, q, NL))) return rc;
    if ((rc = dev_alloc(p, &w.g_num, 2 * NL)) || (rc = dev_alloc(p, &w.g_adfh, DL)) ||
        (rc = dev_alloc(p, &w.eA, 2 * DL)) || (rc = dev_alloc(p, &w.eT, 2 * DL)) ||
        (rc = dev_alloc(p, &w.eud, 4 * DL)) || (rc = dev_alloc(p, &w.g_vf, DL)) ||
        (rc = dev_alloc(p, &w.g_tr, 2 * EL)))
        return rc;
    const std::vector<int>& ud = p->host_updn;
    const std::vector<int>& et = p->host_edge_tri;
    std::vector<int> ud_cnt(m.E + 1, 0), et_cnt(m.E + 1, 0);
    for (int e = 0; e < m.D; ++e) {
        if (ud[2 * e] >= 0 && ud[2 * e + 1] >= 0) {
            ud_cnt[ud[2 * e] + 1]++;
            ud_cnt[ud[2 * e + 1] + 1]++;
        }
        for (int side = 0; side < 2; ++side)
            if (et[2 * e + side] >= 0) et_cnt[et[2 * e + side] + 1]++;
    }
    for (int e = 0; e < m.E; ++e) { ud_cnt[e + 1] += ud_cnt[e]; et_cnt[e + 1] += et_cnt[e]; }
    std::vector<int> ud_ent(ud_cnt[m.E]), et_ent(et_cnt[m.E]);
    std::vector<int> up(ud_cnt.begin(), ud_cnt.end() - 1), ep(et_cnt.begin(), et_cnt.end() - 1);
    for (int e = 0; e < m.D; ++e) {
        if (ud[2 * e] >= 0 && ud[2 * e + 1] >= 0) {
            ud_ent[up[ud[2 * e]]++] = 2 * e;
            ud_ent[up[ud[2 * e + 1]]++] = 2 * e + 1;
        }
        for (int side = 0; side < 2; ++side) {
            const int el = et[2 * e + side];
            if (el >= 0) et_ent[ep[el]++] = 2 * e + side;
        }
    }
    if ((rc = dev_upload(p, &w.ud_off, ud_cnt.data(), ud_cnt.size())) ||
        (rc = dev_upload(p, &w.ud_ent, ud_ent.data(), ud_ent.size())) ||
        (rc = dev_upload(p, &w.et_off, et_cnt.data(), et_cnt.size())) ||
        (rc = dev_upload(p, &w.et_ent, et_ent.data(), et_ent.size())))
        return rc;
    p->adj_ready = true;
    return FCT_OK;
}

int apply_tuning(FusedTuning& t, int knob, int value) {
    switch (knob) {
        case FCT_TUNE_XCD:
            if (value < 1 || value > 64) return set_error(FCT_E_INVALID, "xcd %d out of range", value);
            t.xcd = value;
            return FCT_OK;
        case FCT_TUNE_PAIR_TRACERS:
            if (value != 0 && value != 1)
                return set_error(FCT_E_INVALID, "pair tracers %d not in {0, 1}", value);
            t.pair = value;
            return FCT_OK;
        case FCT_TUNE_W_EDGE_FLUX:
            if (value != 0 && value != 1)
                return set_error(FCT_E_INVALID, "compute_w form %d not in {0, 1}", value);
            t.w_vflux = value;
            return FCT_OK;
        case FCT_TUNE_XCD_CHUNK:
            if (value < 0 || value > (1 << 20))
                return set_error(FCT_E_INVALID, "xcd chunk %d out of range", value);
            t.xchunk = value;
            return FCT_OK;
        case FCT_TUNE_WAVES_PER_GROUP:
            if (value != 1 && value != 2 && value != 4)
                return set_error(FCT_E_INVALID, "waves per workgroup %d not in {1, 2, 4}", value);
            t.wpb = value;
            return FCT_OK;
        case FCT_TUNE_VFLUX_EDGES:
            if (value != 2 && value != 4)
                return set_error(FCT_E_INVALID, "vflux edges per wave %d not in {2, 4}", value);
            t.vflux_epw = value;
            return FCT_OK;
        case FCT_TUNE_ADFH_CHUNK:
            if (value != 0 && value != 4 && value != 8)
                return set_error(FCT_E_INVALID, "adf_h level chunk %d not in {0, 4, 8}", value);
            t.adfh_kc = value;
            return FCT_OK;
        case FCT_TUNE_ADFH_THREADS:
            if (value != 1 && value != 2 && value != 4)
                return set_error(FCT_E_INVALID, "adf_h threads per edge %d not in {1, 2, 4}", value);
            t.adfh_tpe = value;
            return FCT_OK;
        case FCT_TUNE_ADFH_EDGES:
            if (value != 1 && value != 2)
                return set_error(FCT_E_INVALID, "adf_h edges per wave %d not in {1, 2}", value);
            t.adfh_epw = value;
            return FCT_OK;
        default:
            return set_error(FCT_E_INVALID, "unknown tuning knob %d", knob);
    }
}

}  // namespace

extern "C" {

int fct_abi_version(void) { return FCT_ABI_VERSION; }

const char* fct_last_error(void) { return g_last_error.c_str(); }

int fct_plan_create(const fct_mesh_desc* d, fct_plan** out) {
    if (!out) return set_error(FCT_E_INVALID, "out is NULL");
    *out = nullptr;
    int rc = validate_mesh(d);
    if (rc) return rc;
    fct_plan* p = new (std::nothrow) fct_plan();
    if (!p) return set_error(FCT_E_NOMEM, "host allocation failed");
    if (hipGetDevice(&p->device) != hipSuccess) {
        delete p;
        return set_error(FCT_E_HIP, "hipGetDevice failed");
    }
    const int N = d->nod2D, E = d->elem2D, D = d->edge2D, L = d->nl;
    DevMesh& m = p->m;
    m.N = N; m.E = E; m.D = D; m.L = L;

    // --- edges ---
    std::vector<int2> h_edges(D), h_et(D), h_updn(D), h_rng(D);
    std::vector<double4> h_cross(D);
    std::vector<double2> h_ec(D);
    for (int e = 0; e < D; ++e) {
        const int n1 = d->edges[2 * e], n2 = d->edges[2 * e + 1];
        const int el1 = d->edge_tri[2 * e], el2 = d->edge_tri[2 * e + 1];
        const int up = d->edge_up_dn_tri[2 * e], dn = d->edge_up_dn_tri[2 * e + 1];
        h_edges[e] = make_int2(n1, n2);
        h_et[e] = make_int2(el1, el2);
        h_updn[e] = make_int2(up >= 0 ? up : 0, dn >= 0 ? dn : 0);
        // fill_up_dn_grad (tracer_adv.py:351-363): interior edges use the up/dn
        // gradients on [max(ulevels_nod2D_max)-1, min(nlevels_nod2D_min)-1)
        if (up >= 0 && dn >= 0) {
            const int lo = std::max(d->ulevels_nod2D_max[n1], d->ulevels_nod2D_max[n2]) - 1;
            const int hi = std::min(d->nlevels_nod2D_min[n1], d->nlevels_nod2D_min[n2]) - 1;
            h_rng[e] = make_int2(lo, hi);
        } else {
            h_rng[e] = make_int2(0, 0);
        }
        const double* c = d->edge_cross_dxdy + 4 * (size_t)e;
        h_cross[e] = make_double4(c[0], c[1], c[2], c[3]);
        // a = R cos(el1), averaged with el2 when present (tracer_adv.py:211-218)
        double a = R_EARTH * d->elem_cos[el1 >= 0 ? el1 : 0];
        if (el2 >= 0) a = 0.5 * (a + R_EARTH * d->elem_cos[el2]);
        h_ec[e] = make_double2(d->edge_dxdy[2 * e] * a, d->edge_dxdy[2 * e + 1] * R_EARTH);
    }
    // wet level range of each edge: union of its triangles' layer ranges
    std::vector<int2> h_ekr(D);
    for (int e = 0; e < D; ++e) {
        int lo = L, hi = 0;
        for (int side = 0; side < 2; ++side) {
            const int el = d->edge_tri[2 * e + side];
            if (el < 0) continue;
            lo = std::min(lo, d->ulevels[el] - 1);
            hi = std::max(hi, d->nlevels[el] - 1);
        }
        h_ekr[e] = lo < hi ? make_int2(lo, hi) : make_int2(0, 0);
    }
    // --- elements ---
    std::vector<int2> h_elev(E);
    for (int e = 0; e < E; ++e) h_elev[e] = make_int2(d->ulevels[e] - 1, d->nlevels[e] - 1);
    // --- nodes ---
    std::vector<int2> h_nlev(N);
    for (int n = 0; n < N; ++n)
        h_nlev[n] = make_int2(d->ulevels_nod2D[n] - 1, d->nlevels_nod2D[n] - 1);
    // node -> (edge, side), increasing flattened index 2*e+side (the order of the
    // reference's segment sum over stack([f, -f], axis=1), ops.py:56-81)
    std::vector<int> ecnt(N + 1, 0), h_eoff(N + 1, 0), h_ecsr(2 * (size_t)D);
    for (int e = 0; e < D; ++e) { ecnt[h_edges[e].x]++; ecnt[h_edges[e].y]++; }
    for (int n = 0; n < N; ++n) h_eoff[n + 1] = h_eoff[n] + ecnt[n];
    {
        std::vector<int> pos(h_eoff.begin(), h_eoff.end() - 1);
        for (int e = 0; e < D; ++e) {
            h_ecsr[pos[h_edges[e].x]++] = 2 * e;
            h_ecsr[pos[h_edges[e].y]++] = 2 * e + 1;
        }
    }
    // node -> elements, increasing element id (the order of the segment sums over
    // elem_nodes, tracer_adv.py:339-340, 404-405)
    std::vector<int> ncnt(N, 0), h_noff(N + 1, 0), h_ncsr(3 * (size_t)E);
    for (int e = 0; e < E; ++e)
        for (int j = 0; j < 3; ++j) ncnt[d->elem_nodes[3 * e + j]]++;
    for (int n = 0; n < N; ++n) h_noff[n + 1] = h_noff[n] + ncnt[n];
    {
        std::vector<int> pos(h_noff.begin(), h_noff.end() - 1);
        for (int e = 0; e < E; ++e)
            for (int j = 0; j < 3; ++j) h_ncsr[pos[d->elem_nodes[3 * e + j]]++] = e;
    }
    // per-node stencil table of the fused path: edges in the order above with their
    // other end, elements in increasing id with their vertices
    int maxdeg = 0;
    for (int n = 0; n < N; ++n)
        maxdeg = std::max(maxdeg, std::max(h_eoff[n + 1] - h_eoff[n], h_noff[n + 1] - h_noff[n]));
    int dm = maxdeg <= 6 ? 6 : (maxdeg <= 8 ? 8 : 0);
    if (dm && (L > 128 || !fused_mesh_ok(d, h_eoff, h_ecsr, h_noff, h_ncsr))) dm = 0;
    std::vector<int4> h_thdr;
    std::vector<int4> h_tedge, h_telem;
    bool vertex_missing = false;   // an element vertex that is not an edge neighbour
    if (dm) {
        // levels at which the adf_h pass reads a node's averaged gradient gnode: the wet
        // levels of an incident edge outside its up/dn gradient range (all of them where
        // that range is empty), as one range [g_lo, g_hi) per node
        std::vector<int> g_lo(N, L), g_hi(N, 0);
        for (int e = 0; e < D; ++e) {
            const int2 kr = h_ekr[e], ur = h_rng[e];
            if (kr.x >= kr.y) continue;
            const int tl = std::max(kr.x, ur.x), th = std::min(kr.y, ur.y);
            int lo = kr.x, hi = kr.y;
            if (tl < th) {
                const bool above = tl > kr.x, below = th < kr.y;
                if (!above && !below) continue;
                lo = above ? kr.x : th;
                hi = below ? kr.y : tl;
            }
            for (int side = 0; side < 2; ++side) {
                const int v = side ? h_edges[e].y : h_edges[e].x;
                g_lo[v] = std::min(g_lo[v], lo);
                g_hi[v] = std::max(g_hi[v], hi);
            }
        }
        h_thdr.resize(N);
        h_tedge.assign((size_t)N * dm, make_int4(0, 0, 0, 0));
        h_telem.assign((size_t)N * dm, make_int4(0, 0, 0, 0));
        for (int n = 0; n < N; ++n) {
            const int de = h_eoff[n + 1] - h_eoff[n], ne = h_noff[n + 1] - h_noff[n];
            const int gl = g_lo[n] < g_hi[n] ? g_lo[n] : 0, gh = g_lo[n] < g_hi[n] ? g_hi[n] : 0;
            h_thdr[n] = make_int4(d->ulevels_nod2D[n] - 1, d->nlevels_nod2D[n] - 1,
                                  de | (gl << 8) | (gh << 16), ne);
            for (int q = 0; q < de; ++q) {
                const int ent = h_ecsr[h_eoff[n] + q];
                const int e = ent >> 1;
                const int2 nd = h_edges[e];
                // layer ranges of the edge's triangles, one byte each (lo1, hi1, lo2, hi2);
                // a missing triangle gets the empty range [0, 0)
                unsigned packed = 0;
                for (int side = 0; side < 2; ++side) {
                    const int el = d->edge_tri[2 * e + side];
                    if (el < 0) continue;
                    const unsigned lo = (unsigned)(d->ulevels[el] - 1);
                    const unsigned hi = (unsigned)std::max(d->nlevels[el] - 1, d->ulevels[el] - 1);
                    packed |= (lo | (hi << 8)) << (16 * side);
                }
                h_tedge[(size_t)n * dm + q].x = ent;
                h_tedge[(size_t)n * dm + q].y = (ent & 1) ? nd.x : nd.y;
                h_tedge[(size_t)n * dm + q].z = (int)packed;
            }
            for (int q = 0; q < ne; ++q) {
                const int el = h_ncsr[h_noff[n] + q];
                // layer range of element slot q, lo|hi bytes; above them, 4 bits per
                // vertex: the edge slot whose far end it is, or dm for the node itself
                const unsigned lo = (unsigned)(d->ulevels[el] - 1);
                const unsigned hi = (unsigned)std::max(d->nlevels[el] - 1, d->ulevels[el] - 1);
                unsigned codes = 0;
                for (int i = 0; i < 3; ++i) {
                    const int v = d->elem_nodes[3 * el + i];
                    unsigned c = (unsigned)dm;
                    if (v != n) {
                        c = 0xf;
                        for (int q2 = 0; q2 < de; ++q2)
                            if (h_tedge[(size_t)n * dm + q2].y == v) { c = (unsigned)q2; break; }
                        if (c == 0xf) vertex_missing = true;
                    }
                    codes |= c << (4 * i);
                }
                h_tedge[(size_t)n * dm + q].w = (int)(lo | (hi << 8) | (codes << 16));
                h_telem[(size_t)n * dm + q] = make_int4(el, d->elem_nodes[3 * el],
                                                        d->elem_nodes[3 * el + 1],
                                                        d->elem_nodes[3 * el + 2]);
            }
            // padding slots repeat slot 0 (the element-slot range .w stays as set)
            for (int q = de; q < dm && de > 0; ++q) {
                int4& t = h_tedge[(size_t)n * dm + q];
                const int4 f = h_tedge[(size_t)n * dm];
                t.x = f.x; t.y = f.y; t.z = f.z;
            }
            for (int q = ne; q < dm && ne > 0; ++q)
                h_telem[(size_t)n * dm + q] = h_telem[(size_t)n * dm];
        }
        if (vertex_missing) dm = 0;   // the fused kernels read element vertices by slot
    }
    // fused-path edge tables: packed triangle layer ranges and gradient-triangle vertices
    std::vector<int> h_etr;
    std::vector<int4> h_udv;
    bool updn_off_edge = false;   // an up/dn triangle without the edge's node
    if (dm) {
        h_etr.assign(D, 0);
        h_udv.assign(2 * (size_t)D, make_int4(0, 0, 0, 0));
        for (int e = 0; e < D; ++e) {
            unsigned packed = 0;
            for (int side = 0; side < 2; ++side) {
                const int el = d->edge_tri[2 * e + side];
                if (el < 0) continue;
                const unsigned lo = (unsigned)(d->ulevels[el] - 1);
                const unsigned hi = (unsigned)std::max(d->nlevels[el] - 1, d->ulevels[el] - 1);
                packed |= (lo | (hi << 8)) << (16 * side);
            }
            h_etr[e] = (int)packed;
            // up triangle: holds the edge's first node at slot cu; down triangle: its
            // second node at slot cd (the triangles upwind of each end, tracer_adv.py:
            // 351-363); the kernel reads those two vertices from the edge's own columns
            const int u = h_updn[e].x, w = h_updn[e].y;
            const int* eu = d->elem_nodes + 3 * (size_t)u;
            const int* ew = d->elem_nodes + 3 * (size_t)w;
            int cu = 0, cd = 0;
            if (d->edge_up_dn_tri[2 * e] >= 0 && d->edge_up_dn_tri[2 * e + 1] >= 0) {
                cu = cd = -1;
                for (int i = 2; i >= 0; --i) {
                    if (eu[i] == h_edges[e].x) cu = i;
                    if (ew[i] == h_edges[e].y) cd = i;
                }
                if (cu < 0 || cd < 0) { updn_off_edge = true; cu = cd = 0; }
            }
            int ou[2], od[2];
            for (int i = 0, a = 0, b = 0; i < 3; ++i) {
                if (i != cu) ou[a++] = eu[i];
                if (i != cd) od[b++] = ew[i];
            }
            h_udv[2 * (size_t)e] = make_int4(ou[0], ou[1], od[0], od[1]);
            h_udv[2 * (size_t)e + 1] = make_int4(cu | (cd << 2), 0, 0, 0);
        }
        if (updn_off_edge) dm = 0;   // the fused adf_h pass relies on it
    }
    // static QR4C stencil depths (_z_stencil, tracer_adv.py:228-238)
    std::vector<double> Zp(L), h_Zs(4 * (size_t)L);
    for (int k = 0; k < L; ++k) Zp[k] = d->Z[k < L - 1 ? k : L - 2];
    for (int k = 0; k < L; ++k) {
        h_Zs[k] = Zp[k];
        h_Zs[L + k] = Zp[k > 0 ? k - 1 : 0];
        h_Zs[2 * L + k] = k < 2 ? Zp[k] : Zp[k - 2];
        h_Zs[3 * L + k] = Zp[k + 1 < L ? k + 1 : L - 1];
    }

    const size_t NL = (size_t)N * L, DL = (size_t)D * L, EL = (size_t)E * L;
    std::vector<double2> h_aav(NL), h_zzb(L);
    for (size_t i = 0; i < NL; ++i) h_aav[i] = make_double2(d->area[i], d->areasvol[i]);
    for (int k = 0; k < L; ++k) h_zzb[k] = make_double2(h_Zs[k], d->zbar[k]);
    if ((rc = dev_upload(p, &m.edges, h_edges.data(), D)) ||
        (rc = dev_upload(p, &m.edge_tri, h_et.data(), D)) ||
        (rc = dev_upload(p, &m.updn, h_updn.data(), D)) ||
        (rc = dev_upload(p, &m.utri_rng, h_rng.data(), D)) ||
        (rc = dev_upload(p, &m.cross, h_cross.data(), D)) ||
        (rc = dev_upload(p, &m.edge_c, h_ec.data(), D)) ||
        (rc = dev_upload(p, &m.en, d->elem_nodes, 3 * (size_t)E)) ||
        (rc = dev_upload(p, &m.gs, d->gradient_sca, 6 * (size_t)E)) ||
        (rc = dev_upload(p, &m.earea, d->elem_area, (size_t)E)) ||
        (rc = dev_upload(p, &m.elev, h_elev.data(), E)) ||
        (rc = dev_upload(p, &m.nlev, h_nlev.data(), N)) ||
        (rc = dev_upload(p, &m.ecsr_off, h_eoff.data(), (size_t)N + 1)) ||
        (rc = dev_upload(p, &m.ecsr, h_ecsr.data(), h_ecsr.size())) ||
        (rc = dev_upload(p, &m.ncsr_off, h_noff.data(), (size_t)N + 1)) ||
        (rc = dev_upload(p, &m.ncsr, h_ncsr.data(), h_ncsr.size())) ||
        (rc = dev_upload(p, &m.area, d->area, NL)) ||
        (rc = dev_upload(p, &m.areasvol, d->areasvol, NL)) ||
        (rc = dev_upload(p, &m.Zs, h_Zs.data(), h_Zs.size())) ||
        (rc = dev_upload(p, &m.zbar, d->zbar, (size_t)L)) ||
        (rc = dev_upload(p, &m.aav, h_aav.data(), NL)) ||
        (rc = dev_upload(p, &m.zzb, h_zzb.data(), (size_t)L)) ||
        (rc = dev_upload(p, &p->ekr, h_ekr.data(), D))) {
        free_plan(p);
        return rc;
    }
    p->kw = std::max(1, std::min((int)FCT_KCHUNK, d->max_tracers > 0 ? d->max_tracers : 2));
    p->host_edge_tri.assign(d->edge_tri, d->edge_tri + 2 * (size_t)D);
    p->host_updn.assign(d->edge_up_dn_tri, d->edge_up_dn_tri + 2 * (size_t)D);
    p->tab.dm = dm;
    // edge patches of the adf_h patch kernel
    LocalityHost loc;
    if (dm) {
        std::vector<unsigned char> has_updn(D);
        for (int e = 0; e < D; ++e)
            has_updn[e] = d->edge_up_dn_tri[2 * e] >= 0 && d->edge_up_dn_tri[2 * e + 1] >= 0;
        if ((rc = build_locality(N, D, h_edges.data(), h_eoff.data(), h_ecsr.data(), h_udv.data(),
                                 has_updn.data(), h_ekr.data(), h_rng.data(), loc))) {
            free_plan(p);
            return rc;
        }
        p->tab.ep.np = (int)loc.ptab.size();
    }
    // Is the caller's numbering local? On a mesh numbered along the mesh (space-filling
    // or subdivision order, RCM) nearly every edge joins nodes close in number; on a
    // scattered one nearly none does. A local numbering is visited in id order (faster
    // for the node passes and the vertical-flux kernel on such meshes); a scattered one
    // in the breadth-first order of the edge patches, so that the columns gathered at
    // the same time are neighbours on the mesh whatever their addresses.
    // "close": within N / 64 numbers, between 64 and 1024 (a subdivision-ordered sphere
    // has 85-96 % of its edges close, the same meshes numbered at random 3-7 %)
    const int window = std::max(64, std::min(1024, N / 64));
    int64_t close = 0;
    for (int e = 0; e < D; ++e) close += std::abs(h_edges[e].x - h_edges[e].y) <= window;
    const bool local_ids = 2 * close >= (int64_t)D;
    std::vector<int> visit(N);
    if (dm && !local_ids) visit = loc.node_order;
    else for (int n = 0; n < N; ++n) visit[n] = n;
    // node passes: for nl > 64 shallow columns, then deep ones; for nl <= 64 every column
    // is shallow (the list lets two tracers share a wavefront); each set in visit order
    std::vector<int> h_nlist;
    if (dm && L > 64) {
        for (int n : visit)
            if (d->nlevels_nod2D[n] - 1 < 64) h_nlist.push_back(n);
        p->tab.nshallow = (int)h_nlist.size();
        for (int n : visit)
            if (d->nlevels_nod2D[n] - 1 >= 64) h_nlist.push_back(n);
    } else if (dm == 6 || (dm && !local_ids)) {
        h_nlist = visit;
        p->tab.nshallow = N;
    }
    if (dm && D && ((rc = dev_upload(p, &p->tab.ep.ptab, loc.ptab.data(), loc.ptab.size())) ||
                    (rc = dev_upload(p, &p->tab.ep.pnodes, loc.pnodes.data(), loc.pnodes.size())) ||
                    (rc = dev_upload(p, &p->tab.ep.elist, loc.elist.data(), loc.elist.size())) ||
                    (rc = dev_upload(p, &p->tab.ep.eslot, loc.eslot.data(), loc.eslot.size())))) {
        free_plan(p);
        return rc;
    }
    if (dm && D && !local_ids) p->tab.vorder = p->tab.ep.elist;
    if (dm && ((rc = dev_upload(p, &p->tab.hdr, h_thdr.data(), h_thdr.size())) ||
               (!h_nlist.empty() && (rc = dev_upload(p, &p->tab.nlist, h_nlist.data(), h_nlist.size()))) ||
               (rc = dev_upload(p, &m.etr, h_etr.data(), h_etr.size())) ||
               (rc = dev_upload(p, &m.udv, h_udv.data(), h_udv.size())) ||
               (rc = dev_upload(p, &p->tab.edge, h_tedge.data(), h_tedge.size())) ||
               (rc = dev_upload(p, &p->tab.elem, h_telem.data(), h_telem.size())))) {
        free_plan(p);
        return rc;
    }
    // everything allocated so far is the mesh's tables; what follows is the plan's own
    p->tables = new (std::nothrow) TableSet();
    if (!p->tables) {
        free_plan(p);
        return set_error(FCT_E_NOMEM, "host allocation failed");
    }
    p->tables->allocs.swap(p->allocs);
    p->tables->bytes = p->bytes;
    p->bytes = 0;
    if ((rc = alloc_fused_workspace(p))) {
        free_plan(p);
        return rc;
    }
    *out = p;
    return FCT_OK;
}

int fct_plan_clone(const fct_plan* src, int32_t max_tracers, fct_plan** out) {
    if (!out) return set_error(FCT_E_INVALID, "out is NULL");
    *out = nullptr;
    if (!src || !src->tables) return set_error(FCT_E_INVALID, "NULL plan");
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != src->device)
        return set_error(FCT_E_INVALID, "current device %d is not the plan's device %d", cur,
                         src->device);
    fct_plan* p = new (std::nothrow) fct_plan();
    if (!p) return set_error(FCT_E_NOMEM, "host allocation failed");
    p->device = src->device;
    p->m = src->m;
    p->ekr = src->ekr;
    p->tab = src->tab;
    p->tune = src->tune;
    p->kw = std::max(1, std::min((int)FCT_KCHUNK, max_tracers > 0 ? max_tracers : 2));
    p->host_edge_tri = src->host_edge_tri;
    p->host_updn = src->host_updn;
    {
        std::lock_guard<std::mutex> g(table_mutex);
        ++src->tables->refs;
    }
    p->tables = src->tables;
    int rc = alloc_fused_workspace(p);
    if (rc) {
        free_plan(p);
        return rc;
    }
    *out = p;
    return FCT_OK;
}

int fct_plan_get_info(const fct_plan* p, fct_plan_info* info) {
    if (!p || !info) return set_error(FCT_E_INVALID, "NULL argument");
    info->nod2D = p->m.N; info->elem2D = p->m.E; info->edge2D = p->m.D; info->nl = p->m.L;
    info->workspace_bytes = p->bytes + (p->tables ? p->tables->bytes : 0);
    info->table_bytes = p->tables ? p->tables->bytes : 0;
    info->table_id = (uint64_t)(uintptr_t)p->tables;
    info->patch_order = p->tab.vorder != nullptr;
    info->reserved = 0;
    info->device = p->device;
    info->paths = FCT_PATH_BIT_PHASED | (p->tab.dm > 0 ? FCT_PATH_BIT_FUSED : 0);
    return FCT_OK;
}

int fct_plan_set_tuning(fct_plan* p, int32_t knob, int32_t value) {
    if (!p) return set_error(FCT_E_INVALID, "NULL plan");
    return apply_tuning(p->tune, knob, value);
}

int fct_plan_destroy(fct_plan* p) {
    free_plan(p);
    return FCT_OK;
}

}  // extern "C"

namespace {

int check_args(const fct_plan* p, const fct_step_args* a) {
    if (!p || !a) return set_error(FCT_E_INVALID, "NULL plan or args");
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != p->device)
        return set_error(FCT_E_INVALID, "current device %d is not the plan's device %d", cur,
                         p->device);
    if (a->ntracers <= 0 || !a->T || !a->T_old || !a->T_new)
        return set_error(FCT_E_INVALID, "need ntracers > 0 and T, T_old, T_new arrays");
    if (!a->uv || !a->w_e || !a->helem || !a->hnode || !a->hnode_new)
        return set_error(FCT_E_INVALID, "NULL flow or thickness field");
    if ((a->Z3d == nullptr) != (a->zbar3 == nullptr))
        return set_error(FCT_E_INVALID, "Z3d and zbar3 must be given together");
    for (int j = 0; j < a->ntracers; ++j) {
        if (!a->T[j] || !a->T_old[j] || !a->T_new[j])
            return set_error(FCT_E_INVALID, "tracer %d has a NULL field", j);
        if ((const double*)a->T_new[j] == a->T[j] || (const double*)a->T_new[j] == a->T_old[j])
            return set_error(FCT_E_INVALID, "tracer %d: T_new aliases T or T_old", j);
    }
    return FCT_OK;
}

}  // namespace

extern "C" {

static int forward_phased(fct_plan* p, const fct_step_args* a, hipStream_t s) {
    int rc = ensure_phased(p);
    if (rc) return rc;
    const DevMesh& m = p->m;
    if ((rc = launch_vflux(m, a, p->vflux, s))) return rc;
    for (int j = 0; j < a->ntracers; ++j) {
        if ((rc = launch_lo(m, a, j, p->vflux, p->LO, p->tvmax, p->tvmin, s)) ||
            (rc = launch_grad(m, a, j, p->trxy, s)) ||
            (rc = launch_pm(m, a, j, p->vflux, p->LO, p->trxy, p->gnode, p->adfh, p->adfv,
                            p->tvmax, p->tvmin, p->fplus, p->fminus, s)) ||
            (rc = launch_final(m, a, j, p->LO, p->adfh, p->adfv, p->tvmax, p->tvmin,
                               p->fplus, p->fminus, p->pf, p->mf, s)))
            return rc;
    }
    return FCT_OK;
}

int fct_forward_ex(fct_plan* p, const fct_step_args* a, int32_t path, void* stream) {
    int rc = check_args(p, a);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const bool fused_ok = a->w_i == nullptr && p->m.L <= 128 && p->tab.dm > 0;
    if (path == FCT_PATH_FUSED && !fused_ok)
        return set_error(FCT_E_UNSUPPORTED,
                         "fused path needs w_i == NULL, nl <= 128 and node degree <= 8");
    if (path == FCT_PATH_PHASED || (path == FCT_PATH_AUTO && !fused_ok))
        return forward_phased(p, a, s);
    if (path != FCT_PATH_AUTO && path != FCT_PATH_FUSED)
        return set_error(FCT_E_INVALID, "unknown path %d", path);
    return launch_fused(p->m, p->tab, p->ekr, a, p->kw, p->tune, p->vflux, p->fLO, p->fgnode,
                        p->fadfh, p->fpf, p->fadfv, s);
}

int fct_forward(fct_plan* p, const fct_step_args* a, void* stream) {
    return fct_forward_ex(p, a, FCT_PATH_AUTO, stream);
}

int fct_forward_upwind(fct_plan* p, const fct_step_args* a, void* stream) {
    int rc = check_args(p, a);
    if (rc) return rc;
    if (a->w_i || a->Z3d)
        return set_error(FCT_E_INVALID, "the upwind step takes no w_i, Z3d or zbar3");
    if (p->tab.dm == 0)
        return set_error(FCT_E_UNSUPPORTED, "this mesh cannot take the node-table kernels");
    return launch_upwind(p->m, p->tab, p->ekr, a, p->tune, p->vflux, (hipStream_t)stream);
}

int fct_ice_fct_solve(fct_plan* p, const fct_ice_args* a, void* stream) {
    if (!p || !a) return set_error(FCT_E_INVALID, "NULL plan or args");
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != p->device)
        return set_error(FCT_E_INVALID, "current device %d is not the plan's device %d", cur,
                         p->device);
    if (!a->a_ice || !a->m_ice || !a->m_snow || !a->u_ice || !a->v_ice || !a->a_new ||
        !a->m_new || !a->ms_new)
        return set_error(FCT_E_INVALID, "NULL ice field");
    const double* ins[5] = {a->a_ice, a->m_ice, a->m_snow, a->u_ice, a->v_ice};
    const double* outs[3] = {a->a_new, a->m_new, a->ms_new};
    for (const double* o : outs)
        for (const double* i : ins)
            if (o == i) return set_error(FCT_E_INVALID, "an ice output aliases an input");
    if (!p->ice_work) {
        const int E = p->m.E;
        std::vector<double> h(E);
        HIPCHK(hipMemcpy(h.data(), p->m.earea, E * sizeof(double), hipMemcpyDeviceToHost));
        for (int e = 0; e < E; ++e) h[e] = std::sqrt(h[e] / 2.0e8);
        int rc;
        double* sq = nullptr;
        double* tab = nullptr;
        double* work = nullptr;
        if ((rc = dev_alloc(p, &sq, (size_t)E)) ||
            (rc = dev_alloc(p, &tab, ice_table_doubles(E))) ||
            (rc = dev_alloc(p, &work, ice_workspace_doubles(p->m.N))))
            return rc;
        HIPCHK(hipMemcpy(sq, h.data(), (size_t)E * sizeof(double), hipMemcpyHostToDevice));
        // built once, on the null stream, and waited for: later calls may use any stream
        if ((rc = launch_ice_tables(p->m, sq, tab, nullptr))) return rc;
        HIPCHK(hipDeviceSynchronize());
        p->ice_tab = tab;
        p->ice_work = work;
    }
    IceCall c;
    c.dt = a->ice_dt;
    c.diff = a->ice_diff;
    c.gamma = a->ice_gamma_fct;
    c.u = a->u_ice;
    c.v = a->v_ice;
    c.X[0] = a->a_ice; c.X[1] = a->m_ice; c.X[2] = a->m_snow;
    c.out[0] = a->a_new; c.out[1] = a->m_new; c.out[2] = a->ms_new;
    return launch_ice_fct(p->m, c, p->ice_work, p->ice_tab, (hipStream_t)stream);
}

int fct_compute_w(fct_plan* p, const double* uv, const double* helem, double* w,
                  void* stream) {
    if (!p || !uv || !helem || !w) return set_error(FCT_E_INVALID, "NULL plan or field");
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != p->device)
        return set_error(FCT_E_INVALID, "current device %d is not the plan's device %d", cur,
                         p->device);
    const hipStream_t s = (hipStream_t)stream;
    if (p->tune.w_vflux && p->tab.dm && p->m.L <= 128) {
        // edge terms once per edge into the plan's vflux workspace, then summed per node
        int rc = launch_vflux_fused(p->m, p->tab, p->ekr, uv, helem, p->tune, p->vflux, s);
        return rc ? rc : launch_compute_w_vf(p->m, p->tab, p->ekr, p->vflux, w, s);
    }
    return launch_compute_w(p->m, uv, helem, w, s);
}

}  // extern "C"

namespace {

int check_vdiff(const fct_plan* p, const fct_vdiff_args* a) {
    if (!p || !a) return set_error(FCT_E_INVALID, "NULL plan or args");
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != p->device)
        return set_error(FCT_E_INVALID, "current device %d is not the plan's device %d", cur,
                         p->device);
    const int K = a->ntracers;
    if (K < 1 || K > FCT_VDIFF_MAX_TRACERS)
        return set_error(FCT_E_INVALID, "ntracers %d outside 1..%d", K, FCT_VDIFF_MAX_TRACERS);
    if (!a->Kv || !a->hnode_new || !a->T || !a->T_new)
        return set_error(FCT_E_INVALID, "NULL Kv, hnode_new, T or T_new");
    if (a->floor_tracer < -1 || a->floor_tracer >= K)
        return set_error(FCT_E_INVALID, "floor_tracer %d outside -1..%d", a->floor_tracer, K - 1);
    for (int j = 0; j < K; ++j) {
        if (!a->T[j] || !a->T_new[j])
            return set_error(FCT_E_INVALID, "tracer %d has a NULL field", j);
        for (int i = 0; i < K; ++i)
            if (i != j && ((const double*)a->T_new[j] == a->T[i] || a->T_new[j] == a->T_new[i]))
                return set_error(FCT_E_INVALID,
                                 "T_new[%d] aliases tracer %d (only T_new[j] == T[j] is allowed)",
                                 j, i);
    }
    return FCT_OK;
}

// the solves of a checked call, tracers two at a time, on `grid` workgroups whose
// scratch slots start at `scratch`
int run_vdiff(const fct_plan* p, const fct_vdiff_args* a, int grid, double* scratch,
              hipStream_t s) {
    const int K = a->ntracers;
    for (int j0 = 0; j0 < K; j0 += 2) {
        VdiffCall c;
        c.K = std::min(2, K - j0);
        c.dt = a->dt;
        c.Kv = a->Kv;
        c.hnn = a->hnode_new;
        c.Z3d = a->Z3d;
        for (int j = 0; j < c.K; ++j) {
            c.T[j] = a->T[j0 + j];
            c.Tn[j] = a->T_new[j0 + j];
            c.bc[j] = a->bc_surf ? a->bc_surf[j0 + j] : nullptr;
            c.sw[j] = a->sw_3d ? a->sw_3d[j0 + j] : nullptr;
        }
        if (a->floor_tracer >= j0 && a->floor_tracer < j0 + c.K) {
            c.floor_j = a->floor_tracer - j0;
            c.floor_v = a->floor_value;
        }
        const int rc = launch_vdiff(p->m, c, grid, scratch, s);
        if (rc) return rc;
    }
    return FCT_OK;
}

}  // namespace

extern "C" {

int fct_impl_vert_diff(fct_plan* p, const fct_vdiff_args* a, void* stream) {
    int rc = check_vdiff(p, a);
    if (rc) return rc;
    if (!p->vd_scratch) {
        int cus = 0;
        HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->device));
        int grid = 0;
        if ((rc = vdiff_max_grid(cus, &grid))) return rc;
        if ((rc = dev_alloc(p, &p->vd_scratch, vdiff_scratch_doubles(p->m.L, grid)))) return rc;
        p->vd_grid = grid;
    }
    return run_vdiff(p, a, p->vd_grid, p->vd_scratch, (hipStream_t)stream);
}

int fct_impl_vert_diff_static(fct_plan* p, const fct_vdiff_args* a, void* stream) {
    int rc = check_vdiff(p, a);
    if (rc) return rc;
    // workgroups: as many as the fixed scratch has slots for, at most the resident count
    const size_t slot = vdiff_scratch_doubles(p->m.L, 1);
    const size_t slots = (size_t)FCT_VDIFF_STATIC_BYTES / sizeof(double) / slot;
    if (slots < 1)
        return set_error(FCT_E_INVALID,
                         "nl = %d needs %zu bytes of scratch per workgroup, more than the "
                         "%d-byte static scratch", p->m.L, slot * sizeof(double),
                         (int)FCT_VDIFF_STATIC_BYTES);
    int cus = 0;
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->device));
    int grid = 0;
    if ((rc = vdiff_max_grid(cus, &grid))) return rc;
    if ((size_t)grid > slots) grid = (int)slots;
    double* scr = nullptr;
    if ((rc = vdiff_static_scratch(p->device, &scr))) return rc;
    return run_vdiff(p, a, grid, scr, (hipStream_t)stream);
}

int fct_tracer_ab(const double* values, const double* valuesold, double* ttf_ab,
                  int64_t count, void* stream) {
    if (count < 0) return set_error(FCT_E_INVALID, "count %lld < 0", (long long)count);
    if (count && (!values || !valuesold || !ttf_ab))
        return set_error(FCT_E_INVALID, "NULL array");
    return launch_tracer_ab(values, valuesold, ttf_ab, count, (hipStream_t)stream);
}

int fct_stage(fct_plan* p, const fct_stage_args* a, void* stream) {
    if (!p || !a) return set_error(FCT_E_INVALID, "NULL plan or args");
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != p->device)
        return set_error(FCT_E_INVALID, "current device %d is not the plan's device %d", cur,
                         p->device);
    // required inputs and outputs per stage (ADV_FLUX_VER_HO: in[2], in[3] optional)
    static const int need_in[17] = {0, 2, 2, 3, 2, 2, 4, 4, 2, 5, 1, 1, 5, 5, 7, 6, 3};
    static const int need_out[17] = {0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 1, 1};
    const int st = a->stage;
    if (st < 1 || st > 16) return set_error(FCT_E_INVALID, "unknown stage %d", st);
    for (int i = 0; i < need_in[st]; ++i)
        if (!a->in[i]) return set_error(FCT_E_INVALID, "stage %d: input %d is NULL", st, i);
    for (int i = 0; i < need_out[st]; ++i)
        if (!a->out[i]) return set_error(FCT_E_INVALID, "stage %d: output %d is NULL", st, i);
    if (st == FCT_STAGE_ADV_FLUX_VER_HO && ((a->in[2] == nullptr) != (a->in[3] == nullptr)))
        return set_error(FCT_E_INVALID, "Z3d and zbar3 must be given together");
    StageScratch w;
    w.vflux = p->vflux;
    if (st == FCT_STAGE_FILL_UP_DN_GRAD || st == FCT_STAGE_ZALESAK_LIMIT ||
        st == FCT_STAGE_ZALESAK_APPLY || st == FCT_STAGE_ADV_TRA_VERT_IMPL) {
        int rc = ensure_phased(p);
        if (rc) return rc;
        w.gnode = p->gnode; w.tvmax = p->tvmax; w.tvmin = p->tvmin; w.fplus = p->fplus;
        w.fminus = p->fminus; w.pf = p->pf; w.mf = p->mf;
    }
    return launch_stage(p->m, w, st, a->in, a->out, a->dt, (hipStream_t)stream);
}

int fct_forward_phase(fct_plan* p, const fct_step_args* a, int32_t j, int32_t phase,
                      const fct_phase_bufs* b, void* stream) {
    int rc = check_args(p, a);
    if (rc) return rc;
    if (j < 0 || j >= a->ntracers)
        return set_error(FCT_E_INVALID, "tracer index %d out of range", j);
    if (!b || !b->LO || !b->tr_xy || !b->fct_plus || !b->fct_minus)
        return set_error(FCT_E_INVALID, "phase buffers must all be given");
    if ((rc = ensure_phased(p))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const DevMesh& m = p->m;
    switch (phase) {
        case FCT_PHASE_LO:
            if ((rc = launch_vflux(m, a, p->vflux, s))) return rc;
            return launch_lo(m, a, j, p->vflux, b->LO, p->tvmax, p->tvmin, s);
        case FCT_PHASE_GRAD:
            return launch_grad(m, a, j, b->tr_xy, s);
        case FCT_PHASE_PM:
            return launch_pm(m, a, j, p->vflux, b->LO, b->tr_xy, p->gnode, p->adfh, p->adfv,
                             p->tvmax, p->tvmin, b->fct_plus, b->fct_minus, s);
        case FCT_PHASE_FINAL:
            return launch_final(m, a, j, b->LO, p->adfh, p->adfv, p->tvmax, p->tvmin,
                                b->fct_plus, b->fct_minus, p->pf, p->mf, s);
        default:
            return set_error(FCT_E_INVALID, "unknown phase %d", phase);
    }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batches: several plans stepped by one launch per pass and (slots, degree) class
// ---------------------------------------------------------------------------
namespace {

struct BatchSlot {
    bool bound = false;
    std::vector<fct_step_args> args;        // host copies pointing into `ptrs`
    std::vector<std::vector<double*>> ptrs;   // per instance: T, T_old, T_new arrays
    std::vector<int> K;
    std::vector<BatchGroup> groups;
    std::vector<void*> allocs;              // device memory of descriptors and tables
    hipGraphExec_t exec = nullptr;
};

}  // namespace

struct fct_batch {
    int device = -1;
    std::vector<fct_plan*> plans;
    FusedTuning tune;
    BatchSlot slots[FCT_BATCH_SLOTS];
    hipStream_t capture = nullptr;
};

namespace {

void release_slot(BatchSlot& s) {
    // launches already queued may still read the slot's tables
    if (s.exec || !s.allocs.empty()) (void)hipDeviceSynchronize();
    if (s.exec) (void)hipGraphExecDestroy(s.exec);
    s.exec = nullptr;
    for (void* p : s.allocs) (void)hipFree(p);
    s.allocs.clear();
    s.groups.clear();
    s.bound = false;
}

// Copies go through the batch's own stream, so a bind does not wait for compute queued
// on other streams.
template <class T>
int slot_upload(BatchSlot& s, hipStream_t st, const T** out, const std::vector<T>& host) {
    void* d = nullptr;
    const size_t bytes = std::max<size_t>(1, host.size()) * sizeof(T);
    if (hipMalloc(&d, bytes) != hipSuccess)
        return set_error(FCT_E_NOMEM, "hipMalloc(%zu bytes) for a batch table failed", bytes);
    s.allocs.push_back(d);
    if (!host.empty()) {
        HIPCHK(hipMemcpyAsync(d, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    *out = static_cast<const T*>(d);
    return FCT_OK;
}

// per-workgroup (instance, first wavefront) table for one pass kind
void block_table(std::vector<int2>& blk, int inst, int64_t waves, int wpb) {
    for (int64_t w = 0; w < waves; w += wpb) blk.push_back(make_int2(inst, (int)w));
}

int build_slot(fct_batch* b, BatchSlot& s, const fct_step_args* args) {
    const int n = (int)b->plans.size();
    const FusedTuning& t = b->tune;
    const size_t dsz = fused_args_size();
    // (slots, degree) classes in first-appearance order
    std::vector<std::pair<int, int>> keys;
    std::vector<int> cls(n);
    for (int i = 0; i < n; ++i) {
        const fct_plan* p = b->plans[i];
        const std::pair<int, int> key(p->m.L <= 64 ? 1 : 2, p->tab.dm);
        auto it = std::find(keys.begin(), keys.end(), key);
        cls[i] = (int)(it - keys.begin());
        if (it == keys.end()) keys.push_back(key);
    }
    for (size_t c = 0; c < keys.size(); ++c) {
        BatchGroup g;
        g.ns = keys[c].first;
        g.dm = keys[c].second;
        std::vector<unsigned char> desc;
        std::vector<int2> bv, bn, ba, bns, bnd, bpa, bp1, be2, be1;
        const bool split = g.ns == 2;     // every instance of such a group has node lists
        for (int i = 0; i < n; ++i) {
            if (cls[i] != (int)c) continue;
            fct_plan* p = b->plans[i];
            const int inst = g.count++;
            const int K = args[i].ntracers;
            desc.resize(desc.size() + dsz);
            int rc = fill_fused_args(desc.data() + inst * dsz, p->m, p->tab, p->ekr, &args[i], t,
                                     p->vflux, p->fLO, p->fgnode, p->fadfh, p->fpf, p->fadfv);
            if (rc) return rc;
            block_table(bv, inst, vflux_waves(p->m.D, t), t.wpb);
            if (split && !p->tab.nlist)
                return set_error(FCT_E_UNSUPPORTED, "batch: plan without node lists");
            // the single-instance launcher's choices, per instance
            const bool pair = p->tab.nlist && g.dm == 6 && K >= 2 && t.pair;
            if (pair) {
                for (int j0 = 0; j0 < K; j0 += 2)
                    block_table(K - j0 >= 2 ? bpa : bp1, inst | (j0 << 16), p->tab.nshallow, t.wpb);
                if (split) block_table(bnd, inst, node_waves(p->m.N - p->tab.nshallow, K), t.wpb);
            } else if (split) {
                block_table(bns, inst, node_waves(p->tab.nshallow, K), t.wpb);
                block_table(bnd, inst, node_waves(p->m.N - p->tab.nshallow, K), t.wpb);
            } else {
                block_table(bn, inst, node_waves(p->m.N, K), t.wpb);
            }
            if (t.adfh_kc && p->tab.ep.np) {
                for (int j0 = 0; j0 < K; j0 += 2) {
                    std::vector<int2>& be = K - j0 >= 2 ? be2 : be1;
                    for (int q = 0; q < p->tab.ep.np; ++q) be.push_back(make_int2(inst, q | (j0 << 24)));
                }
            } else {
                block_table(ba, inst, adfh_waves(p->m.D, K, t), t.wpb);
            }
        }
        std::vector<unsigned char> dcopy(desc);
        const unsigned char* dd = nullptr;
        int rc;
        if ((rc = slot_upload(s, b->capture, &dd, dcopy)) ||
            (rc = slot_upload(s, b->capture, &g.blk_v, bv)) ||
            (rc = slot_upload(s, b->capture, &g.blk_n, bn)) ||
            (rc = slot_upload(s, b->capture, &g.blk_a, ba)) ||
            (rc = slot_upload(s, b->capture, &g.blk_pa, bpa)) ||
            (rc = slot_upload(s, b->capture, &g.blk_p1, bp1)) ||
            (rc = slot_upload(s, b->capture, &g.blk_e2, be2)) ||
            (rc = slot_upload(s, b->capture, &g.blk_e1, be1)) ||
            (split && ((rc = slot_upload(s, b->capture, &g.blk_ns, bns)) ||
                       (rc = slot_upload(s, b->capture, &g.blk_nd, bnd)))))
            return rc;
        g.desc = reinterpret_cast<const FusedArgs*>(dd);
        g.nb_v = (unsigned)bv.size(); g.nb_n = (unsigned)bn.size(); g.nb_a = (unsigned)ba.size();
        g.nb_ns = (unsigned)bns.size(); g.nb_nd = (unsigned)bnd.size();
        g.nb_pa = (unsigned)bpa.size(); g.nb_p1 = (unsigned)bp1.size();
        g.nb_e2 = (unsigned)be2.size(); g.nb_e1 = (unsigned)be1.size();
        s.groups.push_back(g);
    }
    s.bound = true;
    return FCT_OK;
}

int check_batch(fct_batch* b, int slot) {
    if (!b) return set_error(FCT_E_INVALID, "NULL batch");
    if (slot < 0 || slot >= FCT_BATCH_SLOTS)
        return set_error(FCT_E_INVALID, "batch slot %d out of range", slot);
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != b->device)
        return set_error(FCT_E_INVALID, "current device %d is not the batch's device %d", cur,
                         b->device);
    return FCT_OK;
}

}  // namespace

extern "C" {

int fct_batch_create(fct_plan* const* plans, int32_t count, fct_batch** out) {
    if (!out) return set_error(FCT_E_INVALID, "out is NULL");
    *out = nullptr;
    if (!plans || count <= 0) return set_error(FCT_E_INVALID, "need count > 0 plans");
    fct_batch* b = new (std::nothrow) fct_batch();
    if (!b) return set_error(FCT_E_NOMEM, "host allocation failed");
    for (int i = 0; i < count; ++i) {
        fct_plan* p = plans[i];
        if (!p) { delete b; return set_error(FCT_E_INVALID, "plan %d is NULL", i); }
        if (p->tab.dm == 0) {
            delete b;
            return set_error(FCT_E_UNSUPPORTED, "plan %d cannot take the fused path", i);
        }
        for (int k = 0; k < i; ++k)
            if (plans[k] == p) {
                delete b;
                return set_error(FCT_E_INVALID, "plan %d appears twice (instances share a workspace)", i);
            }
        if (i == 0) b->device = p->device;
        else if (p->device != b->device) {
            delete b;
            return set_error(FCT_E_INVALID, "plans live on different devices");
        }
        b->plans.push_back(p);
    }
    if (hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return set_error(FCT_E_HIP, "hipStreamCreateWithFlags failed");
    }
    *out = b;
    return FCT_OK;
}

int fct_batch_set_tuning(fct_batch* b, int32_t knob, int32_t value) {
    if (!b) return set_error(FCT_E_INVALID, "NULL batch");
    FusedTuning t = b->tune;
    int rc = apply_tuning(t, knob, value);
    if (rc) return rc;
    b->tune = t;
    for (int s = 0; s < FCT_BATCH_SLOTS; ++s) {
        BatchSlot& sl = b->slots[s];
        if (!sl.bound) continue;
        release_slot(sl);
        if ((rc = build_slot(b, sl, sl.args.data()))) return rc;
    }
    return FCT_OK;
}

int fct_batch_bind(fct_batch* b, int32_t slot, const fct_step_args* args) {
    int rc = check_batch(b, slot);
    if (rc) return rc;
    if (!args) return set_error(FCT_E_INVALID, "NULL args");
    const int n = (int)b->plans.size();
    for (int i = 0; i < n; ++i) {
        if ((rc = check_args(b->plans[i], &args[i]))) return rc;
        if (args[i].w_i)
            return set_error(FCT_E_UNSUPPORTED, "instance %d: w_i is not supported in batches", i);
        if (args[i].ntracers > b->plans[i]->kw)
            return set_error(FCT_E_UNSUPPORTED, "instance %d: %d tracers > plan's max_tracers %d",
                             i, args[i].ntracers, b->plans[i]->kw);
    }
    BatchSlot& s = b->slots[slot];
    release_slot(s);
    if ((rc = build_slot(b, s, args))) { release_slot(s); return rc; }
    // keep the arguments (and the caller's pointer arrays' contents) for re-tuning
    s.args.assign(args, args + n);
    s.ptrs.assign(n, {});
    for (int i = 0; i < n; ++i) {
        const int K = args[i].ntracers;
        std::vector<double*>& v = s.ptrs[i];
        v.resize(3 * (size_t)K);
        for (int j = 0; j < K; ++j) {
            v[j] = const_cast<double*>(args[i].T[j]);
            v[K + j] = const_cast<double*>(args[i].T_old[j]);
            v[2 * K + j] = args[i].T_new[j];
        }
        s.args[i].T = v.data();
        s.args[i].T_old = v.data() + K;
        s.args[i].T_new = v.data() + 2 * K;
    }
    return FCT_OK;
}

int fct_batch_launch(fct_batch* b, int32_t slot, void* stream) {
    int rc = check_batch(b, slot);
    if (rc) return rc;
    const BatchSlot& s = b->slots[slot];
    if (!s.bound) return set_error(FCT_E_INVALID, "batch slot %d is not bound", slot);
    for (const BatchGroup& g : s.groups)
        if ((rc = launch_batch_group(g, b->tune, (hipStream_t)stream))) return rc;
    return FCT_OK;
}

int fct_batch_graph_launch(fct_batch* b, int32_t slot, void* stream) {
    int rc = check_batch(b, slot);
    if (rc) return rc;
    BatchSlot& s = b->slots[slot];
    if (!s.bound) return set_error(FCT_E_INVALID, "batch slot %d is not bound", slot);
    if (!s.exec) {
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal));
        for (const BatchGroup& g : s.groups) {
            if ((rc = launch_batch_group(g, b->tune, b->capture))) {
                (void)hipStreamEndCapture(b->capture, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
        }
        HIPCHK(hipStreamEndCapture(b->capture, &graph));
        const hipError_t e = hipGraphInstantiate(&s.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            s.exec = nullptr;
            return set_error(FCT_E_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
        }
    }
    HIPCHK(hipGraphLaunch(s.exec, (hipStream_t)stream));
    return FCT_OK;
}

int fct_batch_destroy(fct_batch* b) {
    if (!b) return FCT_OK;
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur != b->device) (void)hipSetDevice(b->device);
    for (BatchSlot& s : b->slots) release_slot(s);
    if (b->capture) (void)hipStreamDestroy(b->capture);
    if (cur >= 0 && cur != b->device) (void)hipSetDevice(cur);
    delete b;
    return FCT_OK;
}

}  // extern "C"

extern "C" {

int fct_backward(fct_plan* p, const fct_step_args* a, const fct_adjoint_args* g, void* stream) {
    if (!p || !a || !g) return set_error(FCT_E_INVALID, "NULL plan, forward or adjoint args");
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != p->device)
        return set_error(FCT_E_INVALID, "current device %d is not the plan's device %d", cur,
                         p->device);
    if (a->ntracers <= 0 || !a->T || !a->T_old)
        return set_error(FCT_E_INVALID, "need ntracers > 0 and T, T_old arrays");
    if (!a->uv || !a->w_e || !a->helem || !a->hnode || !a->hnode_new)
        return set_error(FCT_E_INVALID, "NULL flow or thickness field");
    if ((a->Z3d == nullptr) != (a->zbar3 == nullptr))
        return set_error(FCT_E_INVALID, "Z3d and zbar3 must be given together");
    if (!g->g_T_new || !g->g_T || !g->g_T_old)
        return set_error(FCT_E_INVALID, "g_T_new, g_T and g_T_old arrays must be given");
    for (int j = 0; j < a->ntracers; ++j)
        if (!a->T[j] || !a->T_old[j] || !g->g_T_new[j] || !g->g_T[j] || !g->g_T_old[j])
            return set_error(FCT_E_INVALID, "tracer %d has a NULL field", j);
    int rc = ensure_adjoint(p);
    if (rc) return rc;
    const size_t NL = (size_t)p->m.N * p->m.L;
    if (a->w_i && !p->aw.LOx &&
        ((rc = dev_alloc(p, &p->aw.LOx, NL)) || (rc = dev_alloc(p, &p->aw.gwi, NL))))
        return rc;
    return launch_backward(p->m, p->aw, a, g->g_T_new, g->g_T, g->g_T_old, g->g_uv, g->g_w_e,
                           g->g_w_i, (hipStream_t)stream);
}

}  // extern "C"
