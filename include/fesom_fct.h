/* C ABI of the MI355X FP64 FCT tracer-advection library (libfesom_fct.so).
 *
 * Drop-in for the reference's FCT entry point
 *   fesom_jax/tracer_adv.py:532-578  advect_one_fct(mesh, uv, w_e, helem, hnode,
 *       hnode_new, T, T_old, *, dt, exch, Z3d, zbar3, w_i) -> (T_new, T_old_new)
 * The reference is pure Python/JAX and has no native FFI; the entry points below are
 * what a ctypes/DLPack binding of that function calls (INTEGRATION.md shows the
 * binding). Plain C types only: host pointers for the mesh, device pointers for
 * fields, a hipStream_t passed as void*.
 *
 * Layout: every 3-D field is row-major [n_entity, nl] float64 (a node's column is
 * contiguous), uv is [elem2D, nl, 2]; indices are 0-based int32 with -1 sentinels
 * (fesom_jax/mesh.py:13-20).
 *
 * Errors: every call returns 0 on success or a negative FCT_E* code; the message of
 * the last failure on the calling thread is returned by fct_last_error().
 * Threading: calls are stream-ordered on the stream passed in. A plan owns one
 * workspace, so two calls on one plan must not overlap in time (serialise them on
 * one stream, or create one plan per stream).
 */
#ifndef FESOM_FCT_H
#define FESOM_FCT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FCT_ABI_VERSION 2

#define FCT_OK 0
#define FCT_E_INVALID (-1)      /* bad argument: null pointer, size, range         */
#define FCT_E_HIP (-2)          /* a HIP runtime call failed                       */
#define FCT_E_NOMEM (-3)        /* device or host allocation failed                */
#define FCT_E_UNSUPPORTED (-4)  /* option the library does not implement           */

typedef struct fct_plan fct_plan;

/* Host-side mesh description: the fields of the reference's Mesh that the FCT reads
 * (fesom_jax/mesh.py:62-129). All pointers are host memory, read only during
 * fct_plan_create. */
typedef struct fct_mesh_desc {
    int32_t nod2D, elem2D, edge2D, nl;
    const int32_t* edges;              /* [edge2D][2] */
    const int32_t* edge_tri;           /* [edge2D][2], -1 = boundary */
    const double*  edge_cross_dxdy;    /* [edge2D][4] metres */
    const double*  edge_dxdy;          /* [edge2D][2] radians */
    const int32_t* edge_up_dn_tri;     /* [edge2D][2], -1 = absent */
    const int32_t* elem_nodes;         /* [elem2D][3] */
    const double*  gradient_sca;       /* [elem2D][6] */
    const double*  elem_area;          /* [elem2D] */
    const double*  elem_cos;           /* [elem2D] */
    const int32_t* nlevels;            /* [elem2D] */
    const int32_t* ulevels;            /* [elem2D] */
    const int32_t* nlevels_nod2D;      /* [nod2D] */
    const int32_t* nlevels_nod2D_min;  /* [nod2D] */
    const int32_t* ulevels_nod2D;      /* [nod2D] */
    const int32_t* ulevels_nod2D_max;  /* [nod2D] */
    const int32_t* nod_in_elem2D_offsets; /* [nod2D+1] */
    const int32_t* nod_in_elem2D;      /* [offsets[nod2D]], element-ascending per node */
    const double*  area;               /* [nod2D][nl] */
    const double*  areasvol;           /* [nod2D][nl] */
    const double*  Z;                  /* [nl-1] mid-layer depths */
    const double*  zbar;               /* [nl] interface depths */
    int32_t max_tracers;               /* tracers per fused launch the workspace holds
                                          (1..8; 0 = 2); more tracers run in chunks */
} fct_mesh_desc;

/* Per-call arguments. Field pointers are device memory of the plan's device.
 * T, T_old, T_new are host arrays of ntracers device pointers. w_i, Z3d, zbar3 may be
 * NULL (the reference's None). T_new[k] must not alias T[k] or T_old[k]. */
typedef struct fct_step_args {
    int32_t ntracers;
    double dt;
    const double* uv;          /* [elem2D][nl][2] */
    const double* w_e;         /* [nod2D][nl] */
    const double* helem;       /* [elem2D][nl] */
    const double* hnode;       /* [nod2D][nl] */
    const double* hnode_new;   /* [nod2D][nl] */
    const double* const* T;    /* ntracers x [nod2D][nl] */
    const double* const* T_old;
    double* const* T_new;
    const double* w_i;         /* optional [nod2D][nl] */
    const double* Z3d;         /* optional [nod2D][nl] (with zbar3) */
    const double* zbar3;       /* optional [nod2D][nl] (with Z3d) */
} fct_step_args;

/* Caller-owned intermediates for the phased path (one tracer). Used when the
 * reference's `exch` hook must see them: LO after phase 1 (tracer_adv.py:559),
 * tr_xy after phase 2 (:562), fct_plus / fct_minus after phase 3 (:438-439). */
typedef struct fct_phase_bufs {
    double* LO;          /* [nod2D][nl] */
    double* tr_xy;       /* [elem2D][nl][2] */
    double* fct_plus;    /* [nod2D][nl] */
    double* fct_minus;   /* [nod2D][nl] */
} fct_phase_bufs;

enum {
    FCT_PHASE_LO = 1,     /* vflux + low-order solution LO (+ implicit w_i solve)    */
    FCT_PHASE_GRAD = 2,   /* element gradients tr_xy                                 */
    FCT_PHASE_PM = 3,     /* node averages, antidiffusive fluxes, bounds, P+/P-      */
    FCT_PHASE_FINAL = 4   /* limiter factors, limited fluxes, T_new                  */
};

enum {
    FCT_PATH_AUTO = 0,    /* fused unless w_i is given or nl > 128                  */
    FCT_PATH_PHASED = 1,  /* one kernel per reference stage                         */
    FCT_PATH_FUSED = 2    /* edge pass + three node-column passes                   */
};

enum {
    FCT_TUNE_XCD = 1,               /* blocks remapped so each of N XCDs sweeps its own
                                       ranges of nodes (1 = hardware order); default 8    */
    FCT_TUNE_WAVES_PER_GROUP = 2,   /* wavefronts per workgroup: 1, 2 or 4 (default 4)     */
    FCT_TUNE_VFLUX_EDGES = 3,       /* edges per wavefront in the vflux pass: 2 or 4 (4)   */
    FCT_TUNE_ADFH_EDGES = 4,        /* edges per wavefront in the adf_h pass: 1 or 2 (2)   */
    FCT_TUNE_XCD_CHUNK = 5,         /* workgroups per range of one XCD under FCT_TUNE_XCD:
                                       0 = one range per XCD, else 1..2^20 (default 8)    */
    FCT_TUNE_W_EDGE_FLUX = 6,       /* fct_compute_w: 1 = edge terms once per edge (the
                                       step's edge flux pass), 0 = per node from the
                                       triangles; same bits either way (default 1)        */
    FCT_TUNE_PAIR_TRACERS = 7,      /* nl > 64, two tracers, node degree <= 6: 1 = the
                                       shallow columns' node passes take both tracers
                                       in one wavefront, 0 = one wavefront per tracer
                                       (default 1)                                         */
    FCT_TUNE_ADFH_CHUNK = 8,        /* adf_h pass: 4 or 8 = edge-patch kernel (edge
                                       patches with their node columns staged in LDS,
                                       that many levels at a time), 0 = one wavefront
                                       per edge column; same bits (default 8)             */
    FCT_TUNE_ADFH_THREADS = 9       /* edge-patch adf_h kernel: threads per edge, 1, 2 or
                                       4, splitting each chunk's levels (default 1)       */
};

#define FCT_BATCH_SLOTS 4

typedef struct fct_plan_info {
    int32_t nod2D, elem2D, edge2D, nl;
    int64_t workspace_bytes;   /* device bytes the plan keeps alive: its own workspace and
                                  its mesh tables (shared with its clones) */
    int32_t device;            /* HIP device ordinal the plan lives on */
    int32_t paths;             /* FCT_PATH_BIT_* of the paths this mesh can take */
    int64_t table_bytes;       /* the mesh tables' share of workspace_bytes */
    uint64_t table_id;         /* equal for plans that share their mesh tables */
    int32_t patch_order;       /* 1: the caller's numbering is not local, so the node
                                  passes and the vertical-flux pass visit nodes and edges
                                  in the edge patches' breadth-first order; 0: id order */
    int32_t reserved;
} fct_plan_info;

#define FCT_PATH_BIT_PHASED 1
#define FCT_PATH_BIT_FUSED 2

/* Library ABI version (FCT_ABI_VERSION). No reference counterpart. */
int fct_abi_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* fct_last_error(void);

/* Build the device-resident plan for a mesh on the current HIP device: topology,
 * CSR node->(edge,side) lists in reference summation order, geometry and workspace.
 * Replaces the per-call mesh handling of fesom_jax/tracer_adv.py:532-578 (the Mesh
 * pytree, fesom_jax/mesh.py:62-129). Allocates device memory: the topology and the
 * fused path's workspace here; the phased path's workspace on the first call that takes
 * that path (exch hook, w_i, or meshes the fused path does not support). */
int fct_plan_create(const fct_mesh_desc* mesh, fct_plan** out);

int fct_plan_get_info(const fct_plan* plan, fct_plan_info* info);

/* A plan for plan's mesh that shares its device tables (reference counted: they are
 * freed with the last plan that uses them) and has its own workspace for max_tracers
 * tracer slots: the cost of a new plan of a known mesh is its workspace only. Results
 * are those of fct_plan_create for the same mesh. No reference counterpart (the
 * reference rebuilds its device arrays per jitted function). */
int fct_plan_clone(const fct_plan* plan, int32_t max_tracers, fct_plan** out);

int fct_plan_destroy(fct_plan* plan);

/* Set a launch-shape knob (FCT_TUNE_*) of the fused path; results do not depend on
 * it. No reference counterpart (XLA chooses its own launch shapes). */
int fct_plan_set_tuning(fct_plan* plan, int32_t knob, int32_t value);

/* One FCT step for args->ntracers tracers sharing the flow inputs:
 * T_new[k] = advect_one_fct(mesh, uv, w_e, helem, hnode, hnode_new, T[k], T_old[k],
 *                           dt=dt, exch=None, Z3d, zbar3, w_i)[0]
 * (fesom_jax/tracer_adv.py:532-578). T_old_new is T itself and is not written.
 * Stream-ordered. On the fused path it allocates nothing; the first call that takes
 * the phased path (w_i given, nl > 128 or node degree > 8) allocates 10 * nod2D * nl +
 * 2 * elem2D * nl + edge2D * nl doubles of workspace, later calls nothing. */
int fct_forward(fct_plan* plan, const fct_step_args* args, void* stream);

/* fct_forward with an explicit kernel path (FCT_PATH_*); both paths give the same
 * bits. No reference counterpart (the reference has one XLA graph). */
int fct_forward_ex(fct_plan* plan, const fct_step_args* args, int32_t path, void* stream);

/* One phase of the step for tracer `tracer` of args (phased path with caller-owned
 * intermediates, so the caller can apply the reference's exch(field, kind) between
 * phases, fesom_jax/tracer_adv.py:559, 562, 438-439). Phases must run in order
 * 1..4 for a tracer; phase 4 writes args->T_new[tracer]. */
int fct_forward_phase(fct_plan* plan, const fct_step_args* args, int32_t tracer,
                      int32_t phase, const fct_phase_bufs* bufs, void* stream);

/* Upwind-only step (the reference's other driver, fesom_jax/tracer_adv.py:154-165):
 * T_new[k] = advect_one(mesh, uv, w_e, helem, hnode, hnode_new, T[k], T_old[k], dt=dt)[0],
 * first-order upwind fluxes of the AB2 field, no limiter. w_i, Z3d and zbar3 must be
 * NULL. Needs a mesh that takes the fused path (FCT_PATH_BIT_FUSED). */
int fct_forward_upwind(fct_plan* plan, const fct_step_args* args, void* stream);

/* ---- sea-ice FCT ---------------------------------------------------------------------
 * (a_new, m_new, ms_new) = fct_solve(cfg, mesh, a_ice, m_ice, m_snow, u_ice, v_ice)
 * (fesom_jax/ice_adv.py:177-194, exch=None): Taylor-Galerkin right-hand side,
 * consistent-mass low- and high-order solves and the element Zalesak limiter of the
 * three ice tracers on the surface nodes. All fields [nod2D], device; outputs must not
 * alias inputs. The first call allocates 32 * nod2D doubles of workspace and 11 doubles
 * per element (the packed element table and its diffusion factor), builds the table and
 * waits for the device to finish it; later calls allocate nothing and do not wait. */
typedef struct fct_ice_args {
    double ice_dt;                    /* IceConfig.ice_dt (fesom_jax/ice.py:62) */
    double ice_diff;                  /* IceConfig.ice_diff (:59) */
    double ice_gamma_fct;             /* IceConfig.ice_gamma_fct (:58) */
    const double* a_ice;
    const double* m_ice;
    const double* m_snow;
    const double* u_ice;
    const double* v_ice;
    double* a_new;
    double* m_new;
    double* ms_new;
} fct_ice_args;

int fct_ice_fct_solve(fct_plan* plan, const fct_ice_args* args, void* stream);

/* ---- vertical velocity ---------------------------------------------------------------
 * w = compute_w(mesh, uv, helem) (fesom_jax/ale.py:309-365), the w_e the step takes:
 * per-level edge divergence into the nodes, cumulative sum from the bottom, divided by
 * the interface area; zero outside each node's interfaces [nzmin, nzmax]. uv is
 * [elem2D][nl][2], helem [elem2D][nl], w [nod2D][nl] (device). nl <= 128. Uses the
 * plan's edge-flux workspace, as fct_forward does: calls on one plan are ordered on one
 * stream (they race if they overlap). No allocation. */
int fct_compute_w(fct_plan* plan, const double* uv, const double* helem, double* w,
                  void* stream);

/* ---- implicit vertical diffusion ---------------------------------------------------
 * The step that follows the FCT in the reference (fesom_jax/step.py:452-456):
 *   T_new[j] = impl_vert_diff_one(mesh, T[j], Kv, hnode_new, dt=dt,
 *                                 bc_surf=bc_surf[j], sw_3d=sw_3d[j], Z3d=Z3d)
 * (fesom_jax/tracer_diff.py:68-128; impl_vert_diff, :131-141, is the two-tracer call
 * with sw_3d on the first tracer only). One tridiagonal solve per node column, the
 * same matrix for every tracer of the call. Optionally clamps one tracer from below on
 * wet layers, the salinity floor of fesom_jax/step.py:456. The first call allocates a
 * scratch bounded by the GPU's resident workgroups (3 * nl * 64 doubles each), later
 * calls allocate nothing. T_new[j] may be T[j] (in place); no other aliasing between
 * the tracers' fields is allowed (FCT_E_INVALID). */
#define FCT_VDIFF_MAX_TRACERS 8
typedef struct fct_vdiff_args {
    int32_t ntracers;                 /* 1 .. FCT_VDIFF_MAX_TRACERS */
    int32_t floor_tracer;             /* tracer clamped to >= floor_value on wet layers; -1: none */
    double dt;
    double floor_value;
    const double* Kv;                 /* [nod2D][nl] interface diffusivity */
    const double* hnode_new;          /* [nod2D][nl] */
    const double* Z3d;                /* [nod2D][nl] layer depths, or NULL: the mesh's Z */
    const double* const* T;           /* ntracers x [nod2D][nl] */
    const double* const* bc_surf;     /* NULL, or ntracers x ([nod2D] or NULL) */
    const double* const* sw_3d;       /* NULL, or ntracers x ([nod2D][nl] or NULL) */
    double* const* T_new;             /* out: ntracers x [nod2D][nl] */
} fct_vdiff_args;

int fct_impl_vert_diff(fct_plan* plan, const fct_vdiff_args* args, void* stream);

/* ---- implicit vertical diffusion that allocates nothing ----------------------------
 * No counterpart in the reference. For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid it): the
 * same arguments, checks and results as fct_impl_vert_diff, bit for bit, but the solver
 * scratch is a fixed device array of FCT_VDIFF_STATIC_BYTES per GPU (reserved when the
 * library is loaded, held until the process ends) instead of plan workspace, and the
 * call makes no allocation. The workgroup count is capped to the array's slots
 * (3 * nl * 64 doubles each), so at large nl the call may run below
 * fct_impl_vert_diff's speed; an nl whose single slot exceeds the array is refused
 * (FCT_E_INVALID). Nothing in the array needs to start at zero. Every call on a GPU
 * shares that GPU's array and leaves its contents behind: calls that overlap on one GPU
 * (other streams, other threads) race. */
#define FCT_VDIFF_STATIC_BYTES (64 << 20)
int fct_impl_vert_diff_static(fct_plan* plan, const fct_vdiff_args* args, void* stream);

/* ---- single stages -----------------------------------------------------------------
 * The reference's public stage functions one at a time, for unit parity
 * (fesom_jax/tracer_adv.py:59-529); fesom_jax_amd.tracer_adv exposes them under the
 * reference's names. All arrays are device pointers in the reference's layouts
 * ([entity][nl], tr_xy [elem2D][nl][2], eud [edge2D][nl][4]); every lane of every output
 * is written, as the reference's whole-array expressions do. Inputs in[] and outputs
 * out[] by stage:
 *   TRACER_AB               in: values, valuesold                    out: ttfAB
 *                           (tracer_ab, :59-66)
 *   EDGE_VFLUX              in: uv, helem                            out: vflux [edge2D][nl] (:69-93)
 *   ADV_FLUX_HOR            in: uv, helem, ttf                       out: flux [edge2D][nl] (:96-105)
 *   ADV_FLUX_VER            in: w_e, ttf                             out: flux [nod2D][nl] (:108-120)
 *   FLUX2DTRACER            in: flux_h, flux_v; dt                   out: del_ttf (:123-139)
 *   ALE_RECONSTRUCT         in: T, del_ttf, hnode, hnode_new         out: T_new (:142-151)
 *   ADV_FLUX_HOR_HO         in: uv, helem, ttf, eud                  out: flux [edge2D][nl] (:196-225)
 *   ADV_FLUX_VER_HO         in: w_e, ttf, Z3d or NULL, zbar3 or NULL out: flux [nod2D][nl] (:241-295)
 *   COMPUTE_FCT_LO          in: flux_h, flux_v, T, hnode, hnode_new; dt   out: LO (:298-313)
 *   TRACER_GRADIENT_ELEMENTS in: ttf                                 out: tr_xy (:316-326)
 *   FILL_UP_DN_GRAD         in: tr_xy                                out: eud (:344-368)
 *   ZALESAK_LIMIT           in: T, LO, adf_h, adf_v, hnode_new; dt   out: adf_h_lim, adf_v_lim
 *                           (:371-466)
 *   ZALESAK_PM              in: as ZALESAK_LIMIT                     out: fct_plus, fct_minus
 *                           (the sums the reference hands to exch, :426-439)
 *   ZALESAK_APPLY           in: as ZALESAK_LIMIT, then fct_plus, fct_minus; dt
 *                           out: adf_h_lim, adf_v_lim (:441-466 on the exchanged sums)
 *   FLUX2DTRACER_FCT        in: T, LO, adf_h, adf_v, hnode, hnode_new; dt  out: del_ttf (:469-484)
 *   ADV_TRA_VERT_IMPL       in: w_i, T, hnode_new; dt                out: T_new (:487-529)
 * The stages share the plan's workspace (edge flux, node-average gradients, limiter
 * sums and factors; the last three groups allocated on first use): calls on one plan
 * must not overlap. */
typedef enum {
    FCT_STAGE_TRACER_AB = 1,
    FCT_STAGE_EDGE_VFLUX = 2,
    FCT_STAGE_ADV_FLUX_HOR = 3,
    FCT_STAGE_ADV_FLUX_VER = 4,
    FCT_STAGE_FLUX2DTRACER = 5,
    FCT_STAGE_ALE_RECONSTRUCT = 6,
    FCT_STAGE_ADV_FLUX_HOR_HO = 7,
    FCT_STAGE_ADV_FLUX_VER_HO = 8,
    FCT_STAGE_COMPUTE_FCT_LO = 9,
    FCT_STAGE_TRACER_GRADIENT_ELEMENTS = 10,
    FCT_STAGE_FILL_UP_DN_GRAD = 11,
    FCT_STAGE_ZALESAK_LIMIT = 12,
    FCT_STAGE_ZALESAK_PM = 13,
    FCT_STAGE_ZALESAK_APPLY = 14,
    FCT_STAGE_FLUX2DTRACER_FCT = 15,
    FCT_STAGE_ADV_TRA_VERT_IMPL = 16
} fct_stage_id;

#define FCT_STAGE_MAX_IN 8
typedef struct fct_stage_args {
    int32_t stage;                     /* fct_stage_id */
    double dt;
    const double* in[FCT_STAGE_MAX_IN];
    double* out[2];
} fct_stage_args;

int fct_stage(fct_plan* plan, const fct_stage_args* args, void* stream);

/* tracer_ab (fesom_jax/tracer_adv.py:59-66) without a plan: ttf_ab[i] =
 * -(0.5+0.1)*valuesold[i] + (1.5+0.1)*values[i] for i < count, on the current device. */
int fct_tracer_ab(const double* values, const double* valuesold, double* ttf_ab,
                  int64_t count, void* stream);

/* ---- reverse mode ------------------------------------------------------------------
 * Cotangents of one FCT step: for the cotangents g_T_new[j] of T_new[j], returns
 * g_T[j], g_T_old[j] and (summed over the tracers) g_uv and g_w_e, with the conventions
 * of jax.grad applied to fesom_jax/tracer_adv.py:532-578 (the reference's backward is
 * jax.grad of the forward, fesom_jax/docs/LIMITER_GRADIENTS.md:45-75: max/min send the
 * cotangent to the winner, half each on a tie; max/min reductions share it among tied
 * arguments; abs'(0) = +1). The forward intermediates are recomputed, not stored.
 * With the implicit split (w_i given) the solve of fesom_jax/tracer_adv.py:555-556 is
 * reversed through the Thomas recurrence (ops.py:121-168) and g_w_i is returned; the
 * moving-level stencil (Z3d, zbar3) is supported. Z3d, zbar3, helem, hnode and
 * hnode_new are constants (no cotangent). The first call allocates the adjoint
 * workspace (and the first with w_i two more [nod2D][nl] arrays); later calls
 * allocate nothing. */
typedef struct fct_adjoint_args {
    const double* const* g_T_new;  /* ntracers x [nod2D][nl], device */
    double* const* g_T;            /* out: ntracers x [nod2D][nl] */
    double* const* g_T_old;        /* out: ntracers x [nod2D][nl] */
    double* g_uv;                  /* out: [elem2D][nl][2] or NULL (not computed) */
    double* g_w_e;                 /* out: [nod2D][nl] or NULL (not computed) */
    double* g_w_i;                 /* out: [nod2D][nl] when fwd->w_i is given, or NULL */
} fct_adjoint_args;

/* fwd: the step's forward arguments (T_new is not read and may be NULL). */
int fct_backward(fct_plan* plan, const fct_step_args* fwd, const fct_adjoint_args* adj,
                 void* stream);

/* ---- batches: several independent instances stepped together -------------------
 * No reference counterpart: the reference steps one mesh per call. A batch groups
 * plans (one per instance, each with its own workspace) and runs one fused FCT step of
 * every instance with one launch per pass and (levels <= 64 or <= 128, node degree)
 * class, over the concatenated task ranges. Results are bit-identical to fct_forward of
 * each instance alone. */
typedef struct fct_batch fct_batch;

/* plans: `count` distinct plans on one device that can take the fused path. */
int fct_batch_create(fct_plan* const* plans, int32_t count, fct_batch** out);

/* Same knobs as fct_plan_set_tuning; rebuilds the bound slots' tables. */
int fct_batch_set_tuning(fct_batch* batch, int32_t knob, int32_t value);

/* Upload the step arguments of every instance (args[count], ntracers <= the plan's
 * max_tracers, w_i == NULL) into slot `slot` (0 <= slot < FCT_BATCH_SLOTS).
 * Synchronous; allocates the slot's device tables. */
int fct_batch_bind(fct_batch* batch, int32_t slot, const fct_step_args* args);

/* One step of every instance with the arguments bound in `slot`, on `stream`. Kernel
 * launches only: no copies, no allocation. */
int fct_batch_launch(fct_batch* batch, int32_t slot, void* stream);

/* As fct_batch_launch, through a HIP graph of the slot's launches that the batch
 * captures on first use (and again after a re-bind or a tuning change). */
int fct_batch_graph_launch(fct_batch* batch, int32_t slot, void* stream);

int fct_batch_destroy(fct_batch* batch);

#ifdef __cplusplus
}
#endif

#endif /* FESOM_FCT_H */
