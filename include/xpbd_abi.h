/*
 * C ABI of the MI355X (gfx950) XPBD Neo-Hookean substep library, libxpbd_hip.so.
 *
 * The reference exposes this path as a Python class, not an FFI:
 *   xpbd_gpu/neohookean_solver.py:15-587  XPBDNeoHookeanSolver
 *   xpbd_gpu/xpbd_base.py:11-204          XPBDSolverBase
 *   xpbd_gpu/sdf.py:29-124                PlaneSdf, SphereSdf, CubeSdf, HangSdfModel
 * Each entry point below names the reference method it replaces.  The Python
 * drop-in package (xpbd_gpu) binds these with ctypes.
 *
 * Conventions
 *  - Every function returns 0 on success or a negative XPBD_E* code; the
 *    message of the last error on the calling thread is read with
 *    xpbd_last_error().
 *  - Host buffers are copied in or out; the library owns all device memory.
 *  - One solver handle is used from one host thread at a time.  All device
 *    work of a handle is ordered on the handle's own HIP stream; the get_*
 *    calls synchronize that stream, the step calls do not.
 *  - Positions and velocities are float32, (n_verts, 3) row-major, in the
 *    caller's vertex order (any internal reordering is hidden).
 */
#ifndef XPBD_ABI_H_
#define XPBD_ABI_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XPBD_OK 0
#define XPBD_EINVAL -1   /* bad argument */
#define XPBD_EHIP -2     /* HIP runtime error */
#define XPBD_ENOMEM -3   /* device allocation failed */
#define XPBD_ESTATE -4   /* call not valid in the current state */
#define XPBD_ENUMERIC -5 /* NaN/Inf found in the state */

#define XPBD_MAX_SDF 8

/* accumulate_mode */
#define XPBD_ACCUM_ATOMIC 0 /* float atomics into per-vertex accumulators */
#define XPBD_ACCUM_CSR 1    /* deterministic vertex-side gather, fixed order */

/* layout: how tets are laid out and processed on the device */
#define XPBD_LAYOUT_FLAT 0  /* one thread per tet, inv_Dm streamed from HBM */
#define XPBD_LAYOUT_PATCH 1 /* one workgroup per patch of <= 512 tets staged in LDS */

/* Internal tet and vertex order (xpbd_desc.reorder; results are returned in
 * the caller's order whatever is chosen):
 *  NONE    the caller's order; patches are cut greedily along it.
 *  BISECT  patch layout: coordinate bisection of the tet centroids into
 *          compact patches, vertices numbered by first appearance; flat
 *          layout: Morton order of tets and vertices.
 *  MORTON  Morton order of tet centroids and vertices; patches are cut
 *          greedily along it. */
#define XPBD_ORDER_NONE 0
#define XPBD_ORDER_BISECT 1
#define XPBD_ORDER_MORTON 2

/* kind of an SDF collider (xpbd_gpu/sdf.py) */
#define XPBD_SDF_PLANE 0  /* PlaneSdf(point, dir, fixed), sdf.py:59-74 */
#define XPBD_SDF_SPHERE 1 /* SphereSdf(center, radius, fixed), sdf.py:77-95 */
#define XPBD_SDF_CUBE 2   /* CubeSdf(center, size, fixed), sdf.py:98-124 */
#define XPBD_SDF_HANG 3   /* HangSdfModel(pos), sdf.py:29-56: union of spheres, always fixed */
#define XPBD_MAX_HANG_POINTS 4096

typedef struct xpbd_sdf {
  int32_t kind;
  int32_t fixed;   /* pins vertices that start inside; always projects (sdf.py:24) */
  double p[4];     /* plane: point xyz | sphere: center xyz, radius | cube: center xyz |
                      hang: p[3] = sphere radius (0.013 in the reference, sdf.py:34) */
  double dir[3];   /* plane: normal (normalized by the library, sdf.py:66) | cube: edge
                      lengths (size; the half size is size * 0.5 in f32, sdf.py:104) */
  const float* points; /* hang: n_points x 3 sphere centres, copied by the library */
  int32_t n_points;    /* hang: 1 .. XPBD_MAX_HANG_POINTS */
  int32_t pad;
  double vel[3];       /* the obstacle's velocity, SdfModel.vel (sdf.py:9); check() treats a
                          vertex as inside when n . (vel - v) > 0 (sdf.py:20-24) */
} xpbd_sdf;

/* Constructor arguments of XPBDNeoHookeanSolver (neohookean_solver.py:18-38)
 * and XPBDSolverBase (xpbd_base.py:15-36), plus the MI355X execution options. */
typedef struct xpbd_desc {
  int32_t n_verts;
  int32_t n_tets;
  const float* rest_positions; /* n_verts*3, host */
  const int32_t* tets;         /* n_tets*4, host, positively oriented */
  double young_modulus;        /* default 2e5 */
  double poisson_ratio;        /* default 0.495 */
  double scale;                /* default 1 */
  double offset[3];            /* default 0 */
  double frame_dt;             /* default 1e-2 */
  double dt;                   /* default 1e-2 */
  int32_t substeps_num;        /* default 20 */
  int32_t constraints_iter;    /* default 5 */
  double hydrostatic_relaxation; /* default 1.2 (split mode) */
  double deviatoric_relaxation;  /* default 1.2 (split mode) */
  int32_t cheb_enable;         /* default 1 */
  int32_t cheb_warmup;         /* default 5 */
  double cheb_gamma;           /* default 0.6666 */
  int32_t block_neohookean_enable; /* default 0 */
  int32_t monitor_convergence; /* default 0: hydro/dev error per iteration */
  double damping;              /* default 5 (xpbd_base.py:35) */
  double gravity[3];           /* default (0,-9.8,0) (xpbd_base.py:33-34) */
  double total_mass;           /* default 1 (xpbd_base.py:36) */
  int32_t n_sdf;               /* >= 1 (xpbd_base.py:66-67) */
  const xpbd_sdf* sdfs;
  /* execution options (no reference counterpart) */
  int32_t accumulate_mode;     /* XPBD_ACCUM_CSR (default) or XPBD_ACCUM_ATOMIC */
  int32_t reorder;             /* XPBD_ORDER_* internal tet/vertex order (default BISECT) */
  int32_t block_size;          /* threads per workgroup for flat tet kernels (64..1024) */
  int32_t device;              /* HIP device ordinal */
  int32_t layout;              /* XPBD_LAYOUT_PATCH (default) or XPBD_LAYOUT_FLAT */
} xpbd_desc;

typedef struct xpbd_solver xpbd_solver;

/* Version of this ABI and the sizes of its structs as compiled, so that a
 * binding can check its own struct layout. */
#define XPBD_ABI_VERSION 5
int xpbd_abi_info(int32_t* version, size_t* desc_size, size_t* sdf_size);

/* Fills *d with the reference defaults (everything but the mesh). */
int xpbd_desc_init(xpbd_desc* d);

/* XPBDNeoHookeanSolver.__init__ (neohookean_solver.py:18-95): uploads the
 * mesh, runs initialize (:97-107) and compute_lumped_mass (xpbd_base.py:133-137). */
int xpbd_create(const xpbd_desc* desc, xpbd_solver** out);
int xpbd_destroy(xpbd_solver* s);

/* Refreshes scalar parameters (material, time step, Chebyshev, mode flags)
 * from desc; the mesh fields of desc are ignored. */
int xpbd_update_params(xpbd_solver* s, const xpbd_desc* desc);

/* self.sdfs = ... (xpbd_base.py:61-70); does not re-pin vertices. */
int xpbd_set_sdfs(xpbd_solver* s, const xpbd_sdf* sdfs, int32_t n);

/* self.gravity[None] = g (xpbd_base.py:33-34) */
int xpbd_set_gravity(xpbd_solver* s, const double g[3]);

/* compute_lumped_mass with self.total_mass = total_mass (xpbd_base.py:89-137) */
int xpbd_compute_lumped_mass(xpbd_solver* s, double total_mass);

/* mesh.verts.x.to_numpy() / from_numpy() (xpbd_base.py:54, utils/renderer.py:49) */
int xpbd_get_positions(xpbd_solver* s, float* xyz);
int xpbd_set_positions(xpbd_solver* s, const float* xyz);
/* Same, with a device pointer (n_verts*3 float32 in device memory, e.g. a
 * torch-ROCm tensor); ordered on the handle's stream, synchronized on return. */
int xpbd_get_positions_device(xpbd_solver* s, void* dev_xyz);
int xpbd_set_positions_device(xpbd_solver* s, const void* dev_xyz);
/* mesh.verts.v, mesh.verts.pred_x, mesh.verts.m / inv_m */
int xpbd_get_velocities(xpbd_solver* s, float* xyz);
int xpbd_set_velocities(xpbd_solver* s, const float* xyz);
int xpbd_get_predicted(xpbd_solver* s, float* xyz);
int xpbd_get_masses(xpbd_solver* s, float* m, float* inv_m);
/* The tets whose volume was at most EPSILON when the masses were last
 * computed (at creation and by xpbd_compute_lumped_mass): their number and
 * the first one in the caller's tet order (-1 when there is none).  The
 * reference prints "Warning: Cell <id> has negative or zero volume" for each
 * of them (xpbd_base.py:104-106) and goes on; so does this library. */
int xpbd_get_degenerate_tets(xpbd_solver* s, int32_t* count, int32_t* first);
/* mesh.cells.lambda_D / lambda_H (neohookean_solver.py:87-92), caller tet order */
int xpbd_get_multipliers(xpbd_solver* s, float* lambda_d, float* lambda_h);
/* indices field (xpbd_base.py:57-58, 78-87): n_tets*12 uint32 face indices */
int xpbd_get_visual_indices(xpbd_solver* s, uint32_t* out);

/* apply_external_forces (xpbd_base.py:139-146) */
int xpbd_apply_external_forces(xpbd_solver* s, double dt);
/* solve_sdf_collision (xpbd_base.py:176-179) */
int xpbd_solve_sdf_collision(xpbd_solver* s, double dt);
/* solve_step (neohookean_solver.py:453-560) */
int xpbd_solve_step(xpbd_solver* s, double dt);
/* advance (xpbd_base.py:148-160) */
int xpbd_advance(xpbd_solver* s, double dt, double damping);
/* apply_external_forces + solve_step + advance (neohookean_solver.py:577-579) */
int xpbd_substep(xpbd_solver* s, double sub_dt);
/* solve() x n_frames (neohookean_solver.py:562-587) */
int xpbd_frame(xpbd_solver* s, int32_t n_frames);

/* Waits for the handle's stream; reports XPBD_EHIP on an asynchronous error. */
int xpbd_synchronize(xpbd_solver* s);
/* The handle's hipStream_t, for callers that order their own work after it. */
int xpbd_get_stream(xpbd_solver* s, void** stream);

/* Per-iteration log.  When enabled, each substep's rows are copied to the
 * host at the end of the substep (one synchronization per substep).
 * Row (8 doubles): frame, substep, iter, r_hat, rho_hat, omega, hydro_error,
 * dev_error.  r_hat is not logged by the reference (neohookean_solver.py:496). */
int xpbd_set_logging(xpbd_solver* s, int32_t enabled);
int xpbd_log_size(xpbd_solver* s, int64_t* rows);
int xpbd_get_log(xpbd_solver* s, double* rows, int64_t max_rows);
int xpbd_clear_log(xpbd_solver* s);

/* Solver counters: current_frame, current_substep, time, rho_hat
 * (neohookean_solver.py:55-56, 65; xpbd_base.py:27). */
int xpbd_get_state(xpbd_solver* s, int32_t* frame, int32_t* substep, double* time,
                   double* rho_hat);
/* Sets the counters at a frame boundary; rho_hat = NaN keeps the device value. */
int xpbd_set_state(xpbd_solver* s, int32_t frame, double time, double rho_hat);

/* Size of the mesh as stored on the device (after padding). */
int xpbd_get_sizes(xpbd_solver* s, int32_t* n_verts, int32_t* n_tets);

/* Execution-path selection (no reference counterpart).
 *  XPBD_PATH_KERNELS:    one launch per kernel per iteration.
 *  XPBD_PATH_GRAPH:      the same kernels, each substep captured in a hipGraph.
 *  XPBD_PATH_PERSISTENT: for small scenes (at most 512 patches of 512 tets,
 *                        the patch layout without isolated vertices, at
 *                        least one iteration), one launch per
 *                        substep that runs every iteration with scene-wide
 *                        waits between its phases; bit-identical to the
 *                        other paths in deterministic mode.  Other scenes take
 *                        the graph path.  If a barrier cannot complete (the
 *                        kernel's workgroups not all resident, e.g. beside
 *                        another persistent kernel), it gives up after 2 s and
 *                        the next synchronizing call reports XPBD_EHIP. */
#define XPBD_PATH_KERNELS 0
#define XPBD_PATH_GRAPH 1
#define XPBD_PATH_PERSISTENT 2
/*  XPBD_PATH_AUTO:       the persistent substep for scenes of at most 288
 *                        patches (where its waits cost less than the phased
 *                        launches' dependent loads), the graph path
 *                        otherwise. */
#define XPBD_PATH_AUTO 3
int xpbd_set_exec_path(xpbd_solver* s, int32_t path);
/* No counterpart in the reference: how long one scene barrier of the
 * persistent substep waits before it gives up (default 2 000 000 us), for
 * callers that share the GPU with other work and for tests of the XPBD_EHIP
 * path (0: give up at once unless every workgroup has already arrived).
 * Microseconds in [0, 1e9]. */
int xpbd_set_barrier_timeout(xpbd_solver* s, int64_t microseconds);

/* Kernel timing (no reference counterpart).  mode 0: off.  mode 1: every
 * launch of the tet projection kernel is timed by HIP events on the solver's
 * stream; in the patch layout the launch records them itself
 * (hipExtLaunchKernel: the kernel's own start and end), in the flat layout
 * they bracket it.  mode r in [2, 1000]: before each projection, r replays of
 * it run back to back, each timed the same way (flat layout: one event pair
 * around all r); lambda is saved before and restored after, so results are
 * unchanged.  xpbd_get_kernel_time returns the summed milliseconds and launch
 * count since the last reset (synchronizes). */
int xpbd_set_kernel_timing(xpbd_solver* s, int32_t enabled);
int xpbd_get_kernel_time(xpbd_solver* s, double* total_ms, int64_t* launches);

/* Reports XPBD_ENUMERIC when any vertex position (x or pred_x) is NaN or
 * infinite; the reference never checks (xpbd_base.py:139-160). */
int xpbd_check_finite(xpbd_solver* s);

/* Device memory held by one solver handle, in bytes (for admission by an
 * HBM budget; no reference counterpart). */
int xpbd_get_device_bytes(xpbd_solver* s, int64_t* bytes);

/* ---- Batched form (no reference counterpart) -------------------------------
 * Many independent scenes advanced together: every kernel launch of a batch
 * substep covers all member scenes (their patches and vertex blocks are
 * concatenated into one work list), and the substep (or a whole frame) is
 * replayed from one hipGraph.  Members must use the patch layout and share
 * constraints_iter, substeps_num, frame_dt, dt and the mode flags; material,
 * mesh, SDFs and masses are per scene.  Results of a scene are bit-identical
 * alone and in any batch (deterministic accumulation).  While a batch runs,
 * member scenes must not be stepped on their own streams;
 * xpbd_batch_synchronize() before reading their state. */
typedef struct xpbd_batch xpbd_batch;

typedef struct xpbd_batch_opts {
  int32_t graph;          /* 1: replay captured graphs (default), 0: launch kernels */
  int32_t graph_substeps; /* substeps per graph: 1 (default) or substeps_num (a frame) */
  int32_t vertex_block;   /* threads per vertex workgroup, 64..1024, multiple of 64 */
  int32_t grid_pct;       /* persistent patch grid, percent of full occupancy, 1..100 */
  int32_t xcd_split;      /* 1: every XCD walks a contiguous share of the work (default) */
  int32_t persistent;     /* the persistent substep (one launch per substep, each member's
                           * workgroups meeting at that member's own barriers):
                           * 0 never; 1 (default) when every member could take it alone,
                           * none spans more than 288 patches and the members hold at
                           * most 512 patches in all (one workgroup per 2 patches); 2 whenever
                           * every member could take it alone and they fit, else the
                           * substep fails with XPBD_EINVAL.  Other batches take the
                           * phased kernels (graph or launches as above). */
} xpbd_batch_opts;

int xpbd_batch_opts_init(xpbd_batch_opts* o);
int xpbd_batch_create(xpbd_solver* const* scenes, int32_t n_scenes, const xpbd_batch_opts* opts,
                      xpbd_batch** out);
int xpbd_batch_destroy(xpbd_batch* b);
/* Re-targets the batch at another member list (same rules as create), reusing
 * its stream and, where large enough, its device buffers; waits for the
 * batch's queued work.  For serving loops whose membership changes at
 * substep boundaries. */
int xpbd_batch_set_members(xpbd_batch* b, xpbd_solver* const* scenes, int32_t n_scenes);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under graph capture, tight serving loops).  Re-targets b like
 * xpbd_batch_set_members, but instead of buffers the batch allocates, its
 * work lists and scene table go to a fixed array of device memory, one per
 * GPU, that the library holds until the process ends; the call makes no
 * device allocation or release.  Limits: at most XPBD_NOALLOC_MAX_SCENES
 * members and XPBD_NOALLOC_LIST_BYTES of work lists (32 B per patch of 512
 * tets, 8 B per vertex block and per persistent workgroup, 16 B per
 * entry of the persistent substep's workgroup list: one per workgroup, or
 * 8 per row when the members are laid out per XCD); larger member
 * lists are refused with XPBD_EINVAL and leave the batch as it was.  The
 * batch must not replay graphs (opts.graph == 0: capturing a new one would
 * allocate).  The array holds one batch's lists per GPU: a later call for
 * another batch on the same GPU takes it over, and the earlier batch then
 * reports XPBD_ESTATE until re-targeted; calls that overlap on one GPU race. */
#define XPBD_NOALLOC_MAX_SCENES 64
#define XPBD_NOALLOC_LIST_BYTES (1 << 20)
int xpbd_batch_set_members_noalloc(xpbd_batch* b, xpbd_solver* const* scenes, int32_t n_scenes);
/* one substep (sub_dt = dt / substeps_num of each scene) of every member */
int xpbd_batch_substep(xpbd_batch* b);
/* n_frames frames of every member (solve(), neohookean_solver.py:562-587) */
int xpbd_batch_frame(xpbd_batch* b, int32_t n_frames);
/* Waits for the batch's queued work; XPBD_EHIP when a persistent substep's
 * barrier gave up since the last call (its results are invalid). */
int xpbd_batch_synchronize(xpbd_batch* b);
/* The path the batch's next substeps take with its current members:
 * XPBD_PATH_PERSISTENT, XPBD_PATH_GRAPH or XPBD_PATH_KERNELS. */
int xpbd_batch_path(xpbd_batch* b, int32_t* path);
/* The batch's xpbd_set_barrier_timeout (no counterpart in the reference). */
int xpbd_batch_set_barrier_timeout(xpbd_batch* b, int64_t microseconds);
int xpbd_batch_get_stream(xpbd_batch* b, void** stream);
int xpbd_batch_info(xpbd_batch* b, int32_t* n_scenes, int64_t* n_verts, int64_t* n_tets,
                    int32_t* n_patches);
/* Persistent grids of the batch's patch kernels (workgroups of the projection
 * and of the residual) and the largest number of scenes one projection
 * workgroup's patch range touches (2 or more: ranges cross scenes). */
int xpbd_batch_grid(xpbd_batch* b, int32_t* grid_project, int32_t* grid_residual,
                    int32_t* max_scenes_per_range);

/* Host-only description of the patch layout a mesh would get (no device
 * work): number of patches, number of (patch, vertex) slots, and the largest
 * vertex and tet counts of a patch.  With positions (n_verts*3) the tets are
 * ordered as the solver orders them (coordinate bisection); with NULL they
 * are cut greedily in the order given.  The layout's invariants are checked
 * and a violation is reported as XPBD_ESTATE. */
int xpbd_patch_stats(const float* positions, const int32_t* tets, int32_t n_tets,
                     int32_t n_verts, int32_t* n_patches, int64_t* n_slots,
                     int32_t* max_patch_verts, int32_t* max_patch_tets);

/* Copies the last error message of the calling thread into buf. */
int xpbd_last_error(char* buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif /* XPBD_ABI_H_ */
