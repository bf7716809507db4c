/*
 * astg.h - C ABI of the MI355X-native pressure projection for the adaptive
 * staggered-tilted (AST) octree.
 *
 * Plain C types only: pointers, sizes, status codes. No exception crosses the
 * boundary; every entry point returns ASTG_OK (0) or a negative ASTG_ERR_*
 * code, and astg_last_error() describes the last failure on a context.
 *
 * Reference interfaces replaced (citations are into the AST-Grid sources):
 *   - the CUDA solver seam `CudaSolver::solve(n, rowstart, colindex, value, b, x)`
 *     with `solver_error` / `solver_iters`
 *     (src/octree_omnigrid_solver3.cpp:664-669, include/octree_omnigrid_solver3.h:101-105)
 *     -> astg_pcg_solve_csr
 *   - `OctreeOmniGridSolver3::project(dt)` (src/octree_omnigrid_solver3.cpp:350-679),
 *     i.e. numbering + assembly + MIC(0)-PCG + `updateVelocityByPressure`
 *     (681-847) -> astg_set_topology_* + astg_set_velocity + astg_project
 *   - `PCGSolver<T>::solve` / `factor_modified_incomplete_cholesky0`
 *     (include/pcgsolver/pcg_solver.h:89-190, 366-426) -> astg_pcg_solve_csr
 *
 * Threading: a context owns one HIP stream and its device buffers; use each
 * context from one host thread at a time. Distinct contexts may be used from
 * different threads: the octree kernels read per-topology tables from
 * constant memory that exists once per GPU, so calls that launch them take a
 * process-wide lock and a change of context re-uploads the tables (after
 * synchronising the GPU).
 */
#ifndef ASTG_H
#define ASTG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ASTG_OK 0
#define ASTG_ERR_ARG (-1)           /* invalid argument or size */
#define ASTG_ERR_HIP (-2)           /* HIP runtime / device error */
#define ASTG_ERR_NOT_CONVERGED (-3) /* PCG hit max_iterations, or rho == 0 / NaN */
#define ASTG_ERR_TOPOLOGY (-4)      /* topology missing or inconsistent */
#define ASTG_ERR_STATE (-5)         /* call order violated (e.g. project before topology) */
#define ASTG_ERR_IO (-6)            /* file missing, truncated or not matching the context */

/* Reference wall flags (include/constants.h:11-26). */
#define ASTG_WALL_XNEG (1 << 0)
#define ASTG_WALL_XPOS (1 << 1)
#define ASTG_WALL_ZNEG (1 << 2)
#define ASTG_WALL_ZPOS (1 << 3)
#define ASTG_WALL_YNEG (1 << 4)
#define ASTG_WALL_YPOS (1 << 5)

/* Precision modes of the projection solve. */
#define ASTG_PREC_FP64 0  /* FP64 MIC(0)-PCG, the reference's arithmetic */
#define ASTG_PREC_MIXED 1 /* FP32 inner PCG, FP64 MIC pivots and FP64 outer residual b-Ax */

/* Ordering of the unknowns in the projection's solve (astg_project,
 * astg_project_fields and the differentiable step; the CSR seam and the
 * batched solves keep the caller's order). The reference numbering gives the
 * reference's MIC(0) factor and iterates; the block-colour ordering (8
 * colours of 8^3-cell cubes, the reference index inside a cube) factors the
 * same matrix in another order: a different preconditioner with far fewer
 * dependency levels in its triangular solves (at 256^3: 67 against 262) and
 * more iterations (353 against 267 to 1e-6 in FP64). The solution, the system
 * astg_get_system returns and the velocity update stay in the reference
 * numbering either way. */
#define ASTG_ORDER_AUTO 0         /* reference numbering for FP64, block colours for the mixed form (the \
                                     block-colour copy is built on the device: 26 ms at 256^3) */
#define ASTG_ORDER_REFERENCE 1
#define ASTG_ORDER_BLOCK_COLOUR 2

/* mic_tau = ASTG_TAU_AUTO: the reference's MIC(0) (tau 0.97) in FP64, where
 * the factor and iterates follow the reference's; plain IC(0) (tau 0) in the
 * mixed form, an added form whose contract is the FP64 residual: on the
 * octree systems here IC(0) needs far fewer iterations than MIC(0) at 0.97
 * (256^3: 176 against 363 in block colours, 171 against 266 in the reference
 * numbering; 10-42 % fewer on the 24 batched scenes). Any tau >= 0 is used as
 * given in either form. */
#define ASTG_TAU_AUTO (-1.0)

/* Norm of the convergence test. */
#define ASTG_NORM_INF 0 /* reference: ||r||_inf <= tol * ||b||_inf (pcg_solver.h:376-408) */
#define ASTG_NORM_L2 1

typedef struct astg_ctx astg_ctx;

typedef struct astg_pcg_params {
    double tol;            /* relative tolerance factor (reference SOLVER_ACCURACY = 1e-5) */
    int max_iterations;    /* reference initData: 3000 */
    int norm;              /* ASTG_NORM_* */
    double mic_tau;        /* MIC(0) modification parameter (reference 0.97); ASTG_TAU_AUTO (the default):
                              0.97 in FP64, 0 (plain IC(0)) in the mixed form */
    double min_diag_ratio; /* small-pivot fallback ratio, reference 0.25 */
    int precision;         /* ASTG_PREC_* */
    int max_outer;         /* mixed form: max refinement sweeps (0 = default 20) */
    int ordering;          /* ASTG_ORDER_* (default ASTG_ORDER_AUTO) */
} astg_pcg_params;

typedef struct astg_pcg_result {
    double residual;     /* final ||r|| in the chosen norm (absolute, like solver_error) */
    double rel_residual; /* residual / ||b|| */
    int iterations;      /* PCG iterations (mixed form: total inner iterations) */
    int outer;           /* refinement sweeps (mixed form), 0 otherwise */
    int converged;       /* 1 if the tolerance was met */
    int n;               /* unknowns */
    long long nnz;       /* stored entries of A */
    int depth;           /* dependency levels of the triangular sweeps */
    double ms_assemble, ms_factor, ms_solve, ms_velocity, ms_total; /* device times */
} astg_pcg_result;

/* Defaults: tol 1e-5, 3000 iterations, inf-norm, tau 0.97, ratio 0.25, FP64. */
void astg_pcg_params_default(astg_pcg_params* p);

int astg_create(astg_ctx** out, int device, uint32_t flags);
int astg_destroy(astg_ctx* ctx);
const char* astg_last_error(const astg_ctx* ctx);

/* ---- CUDA-solver seam -------------------------------------------------- */
/* Solves A x = b with MIC(0)-PCG on the device. A is CSR (rowstart n+1,
 * colindex/value nnz), symmetric, with rows sorted by column as the
 * reference's FixedSparseMatrix (include/pcgsolver/sparse_matrix.h:276-325).
 * Host pointers; x is fully overwritten. Returns ASTG_ERR_NOT_CONVERGED when
 * the tolerance is not met (x still holds the last iterate).
 * Limit: at most 16 entries on each side of the diagonal per row (a sweep row
 * keeps its dependencies in registers); a wider row is rejected with
 * ASTG_ERR_ARG and the message "row i has more than 16 entries on one side of
 * the diagonal (unsupported)". The octree's matrices have at most 15 entries
 * per row, diagonal included, so no matrix project() assembles reaches it.
 * Unsorted rows, out-of-range columns and an unsymmetric pattern are rejected
 * with ASTG_ERR_ARG too. */
int astg_pcg_solve_csr(astg_ctx* ctx, int n, const int* rowstart, const int* colindex, const double* value,
                       const double* b, double* x, const astg_pcg_params* params, astg_pcg_result* result);

/* Batched form: `count` independent systems, each given as to
 * astg_pcg_solve_csr (system k: n[k] rows, rowstart[k], colindex[k],
 * value[k], b[k], output x[k]), solved together. The systems are stacked
 * block-diagonally and factored in one pass; every system keeps its own PCG
 * scalars, convergence test and iteration count, so each follows the
 * recurrence astg_pcg_solve_csr would run on it alone (up to the summation
 * order of its dots). results[k] (may be null) per system; its device times
 * are those of the whole batch. Replaces a loop of CudaSolver::solve calls,
 * one per scene (octree_omnigrid_solver3.cpp:666-669); the reference has no
 * batched solver. Returns ASTG_ERR_NOT_CONVERGED if any system missed the
 * tolerance (every x[k] still holds its last iterate). */
int astg_pcg_solve_batch(astg_ctx* ctx, int count, const int* n, const int* const* rowstart,
                         const int* const* colindex, const double* const* value, const double* const* b,
                         double* const* x, const astg_pcg_params* params, astg_pcg_result* results);

/* Launch shape of batched solves (no counterpart in the reference). */
typedef struct astg_batch_tuning {
    int merge_levels;         /* dependency levels per merged sweep group, 1-8 (default 2) */
    int stream_blocks_per_cu; /* grid of the streaming kernels, blocks per compute unit, 1-64 (default 8) */
    int lanes_per_row;        /* SpMV lanes per row: 4, 8 (default) or 16 */
    double ms_per_iteration;  /* astg_pcg_tune_batch: one iteration of the chosen shape, ms */
} astg_batch_tuning;
void astg_batch_tuning_default(astg_batch_tuning* t);
/* Sets the shape used by the context's next astg_pcg_solve_batch calls. */
int astg_pcg_set_batch_tuning(astg_ctx* ctx, const astg_batch_tuning* t);
/* Autotuner: times `iterations` PCG iterations (params' precision; no system
 * converges) of the given batch for merge_levels in {1,2,3} x blocks per CU
 * in {4,8,16} x lanes per row in {4,8,16}, best of two runs each, then sets
 * the fastest shape on the context and returns it in *best. Results of the
 * solves are discarded. */
int astg_pcg_tune_batch(astg_ctx* ctx, int count, const int* n, const int* const* rowstart,
                        const int* const* colindex, const double* const* value, const double* const* b,
                        const astg_pcg_params* params, int iterations, astg_batch_tuning* best);

/* ---- scene helpers ---------------------------------------------------- */
/* Split map (R^3 ints, x fastest, value = leaf level) refined toward a union
 * of spheres (n x {cx,cy,cz,r}); band[l] = refinement distance for level l+1 -> l;
 * then balanced so every leaf's 26 same-level neighbour blocks exist. */
int astg_split_spheres(int R, int levels, int nspheres, const double* spheres, const double* band, int* split_out);

/* A split map from a distance field instead of spheres: dist holds
 * (R >> dist_shift)^3 values at the centres of cells 2^dist_shift finest cells
 * wide (x fastest; 0 <= dist_shift < levels; e.g. the distance to the region
 * that must be finest, <= 0 inside it); a level-l leaf is split while the
 * smallest dist over its cells is within band[l-1], then the map is balanced
 * as above. Used to rebuild the layout from a density field (no reference
 * counterpart: the reference builds its layout once, initData 68). */
int astg_split_distance(int R, int levels, int dist_shift, const float* dist, const double* band, int* split_out);

/* ---- topology (reference node order) ---------------------------------- */
/* Builds layers, ghosts, T-joint / half-tilt state (constrainOctreeTiltE3,
 * octree_util.cpp:38-169) and initData's fix-ups (octree_omnigrid_solver3.cpp:128-145)
 * from a split map. closed_flags: ASTG_WALL_* bits of closed walls. No collider
 * until astg_set_boundary_sdf (boundary SDF > 0 everywhere, so all cutting-cell
 * weights are 1 except half-tilt slots, initCuttingCellWeights 1169-1238). */
int astg_set_topology_split(astg_ctx* ctx, int R, double h0, int levels, const int* split, int closed_flags);

/* Boundary (collider) SDF, as the reference's initObstacle installs one
 * (octree_omnigrid_solver3.cpp:152-194, omni_boundary_solver3.cpp:11-31): sdf
 * holds R^3 signed distances at the finest cell centres ((i+0.5) h0, x
 * fastest), negative inside the collider; NULL removes the collider. It sets
 * the unknown numbering's -3h cut (updateIndexParallel 292-347), the
 * cutting-cell weights (initCuttingCellWeights 1169-1238), the slip constraint
 * (constrainVelocity, omni_boundary_solver3.h:26-100), the back-trace clip
 * (BackTrace, advection.h:45-62) and astg_wipe_boundary. slip_center (3
 * doubles; NULL keeps the reference's hard-coded (0.5, 0.5, 0.45)) is where
 * the slip normals start. Like astg_set_topology_split, the fields restart at
 * zero; astg_set_topology_split removes the collider. */
int astg_set_boundary_sdf(astg_ctx* ctx, const double* sdf, const double* slip_center);
/* R^3 signed distances |p - c| - radius of a sphere at the finest cell
 * centres (the reference's Sphere3 collider; its own grid agrees to 1 ulp). */
int astg_sphere_sdf(int R, double h0, double cx, double cy, double cz, double radius, double* out);

/* Opens non-fixed level-l leaf tilts where mask[i] != 0 (e = h_l/2), closes the
 * others, as adjustE commits its decision (octree_omnigrid_solver3.cpp:1150-1165). */
int astg_set_tilt_open(astg_ctx* ctx, int level, const unsigned char* mask);

int astg_level_info(astg_ctx* ctx, int level, int* nnode, int* nleaf);
/* xyz: 3*nnode ints; flags: nnode ints (1 valid, 2 ghost, 4 ghost-from-coarse). */
int astg_get_nodes(astg_ctx* ctx, int level, int* xyz, int* flags);
/* bits: 1 largest, 2 closed, 4 fixed, 8 T-joint, 16 closed-by-T-joint, 32 half-tilt,
 * 64 layout open, 128 layout fixed. */
int astg_get_tilt(astg_ctx* ctx, int level, double* e, int* bits, int* half_dir);
/* count receives the number of half-tilts; lxyz (may be NULL) 4*count ints. */
int astg_get_half_tilts(astg_ctx* ctx, int* lxyz, int* count);
/* Unknown numbering per node (-1 = not an unknown); n_total may be NULL. */
int astg_get_numbering(astg_ctx* ctx, int level, int* oct, int* tilt, int* n_total);
/* Cutting-cell weights: mac 3*nnode, tilt 8*nnode. */
int astg_get_weights(astg_ctx* ctx, int level, double* mac, double* tilt8);

/* ---- fields (per level, node order) ------------------------------------ */
/* mac: 3*nnode (x,y,z per node), tilt8: 8*nnode. Host pointers. */
int astg_set_velocity(astg_ctx* ctx, int level, const double* mac, const double* tilt8);
int astg_get_velocity(astg_ctx* ctx, int level, double* mac, double* tilt8);

/* ---- projection ------------------------------------------------------- */
/* One pressure projection on the device: numbering (cached with topology),
 * assembly of A and b, MIC(0) factorisation, PCG, velocity correction.
 * Velocities are read from and written back to the context's device copy. */
int astg_project(astg_ctx* ctx, double dt, double rho, const astg_pcg_params* params, astg_pcg_result* result);

/* Last assembled system: rowstart n+1, col/val nnz, b n (scaled by rho/dt), p n.
 * Any pointer may be NULL. On a host-only context (device < 0) only the
 * pattern (rowstart, col) is available. */
int astg_get_system(astg_ctx* ctx, int* rowstart, int* col, double* val, double* b, double* p);

/* MIC(0) factor of the last assembled system, in the reference's column form
 * (SparseColumnLowerFactor, pcg_solver.h:22-79): invdiag n, colstart n+1,
 * rowindex / value nnz_L. Pass NULL arrays to query nnz_L. */
int astg_get_mic_factor(astg_ctx* ctx, double* invdiag, unsigned* colstart, unsigned* rowindex, double* value,
                        long long* nnz_l);

/* ---- fields in the caller's memory (PyTorch-ROCm and other device users) */
/* The context's field layout: every level's nodes concatenated (level l holds
 * nodes [offset[l], offset[l+1]) in node order, G = offset[levels]). MAC
 * velocity 3*G axis-major (u_x of node g at g, u_y at G+g, u_z at 2G+g), tilt
 * velocity 8*G slot-major, each scalar one cell and one vertex array of G.
 * offsets: levels+1 ints (may be NULL). */
int astg_field_layout(astg_ctx* ctx, long long* G, long long* offsets);

/* Pointers to caller-owned fields in that layout. A NULL array is skipped.
 * device_pointers = 1: device memory of the context's GPU, copied device to
 * device in the order of `stream` (the caller's hipStream_t, e.g. torch's
 * current stream; NULL = the null stream): the copies wait for the work
 * already queued on `stream`, and work queued on `stream` afterwards waits
 * for them; the host is not blocked. device_pointers = 0: host memory,
 * copied synchronously (stream ignored). */
typedef struct astg_fields {
    double* mac;          /* 3*G */
    double* tilt;         /* 8*G */
    double* density_cell; /* G */
    double* density_vert; /* G */
    double* heat_cell;    /* G */
    double* heat_vert;    /* G */
    int device_pointers;
    void* stream;
} astg_fields;
/* Copies the caller's fields into the context (before project / advect). */
int astg_fields_upload(astg_ctx* ctx, const astg_fields* f);
/* Copies the context's current fields out to the caller's arrays. */
int astg_fields_download(astg_ctx* ctx, const astg_fields* f);
/* project(dt) on caller-owned fields (the matrix-free seam of
 * OctreeOmniGridSolver3::project, src/octree_omnigrid_solver3.cpp:350-679,
 * with updateVelocityByPressure 681-847): the velocity in `inout` is copied
 * in, projected and copied back in place, in `inout->stream` order. */
int astg_project_fields(astg_ctx* ctx, double dt, double rho, const astg_pcg_params* params, astg_fields* inout,
                        astg_pcg_result* result);

/* ---- workload replay and diagnostics ---------------------------------- */
/* Saves (restore = 0) or restores (restore = 1) a device-side copy of the
 * velocity fields, so the same input frame can be projected repeatedly
 * without host traffic. No counterpart in the reference. */
int astg_fields_snapshot(astg_ctx* ctx, int restore);

/* Average device time of one kernel of the current projection system over
 * reps launches, timed with HIP events on the context's stream, and the
 * kernel's algorithmic HBM bytes per launch. which: ASTG_K_*. Needs one
 * astg_project call first. No counterpart in the reference. */
#define ASTG_K_SPMV64 0  /* q = A s fused with s.q (FP64) */
#define ASTG_K_FWD64 1   /* forward MIC sweep (FP64) */
#define ASTG_K_BWD64 2   /* backward MIC sweep fused with z.r (FP64) */
#define ASTG_K_MIC 3     /* MIC(0) factorisation */
#define ASTG_K_ASSEMBLE 4 /* assembly of A and b */
#define ASTG_K_SPMV32 5  /* q = A s fused with s.q (FP32 values and vectors) */
#define ASTG_K_FWD32 6   /* forward MIC sweep (FP32) */
#define ASTG_K_BWD32 7   /* backward MIC sweep (FP32) */
#define ASTG_K_COPY 8    /* plain copy of n doubles (HBM counter calibration) */
int astg_kernel_time(astg_ctx* ctx, int which, int reps, double* avg_ms, double* bytes);

/* Per-row timeline of one launch of an FP64 sweep (ASTG_K_FWD64 or
 * ASTG_K_BWD64) of the current system: t_out[3*i + {0,1,2}] = device wall
 * clock (100 MHz ticks) when row i's wavefront passed its gate, when all of
 * row i's dependencies were seen, and after row i published its value;
 * level_out[i] = dependency level of row i. Arrays of 3n and n entries. For
 * latency analysis; no counterpart in the reference. */
int astg_sweep_timeline(astg_ctx* ctx, int which, long long* t_out, int* level_out);

/* ---- the step around the projection and advection --------------------- */
/* Scalar fields carried by the context: ASTG_FIELD_DENSITY, ASTG_FIELD_HEAT;
 * per level, one octagon (cell) and one tilt (vertex) value per node, in node
 * order, as OctreeOmniGrid3<REAL> stores them (include/octree_omni_grid3.h).
 * Scalars and velocity start at zero and are reset when the topology is set
 * again (astg_set_topology_split, astg_set_tilt_open); tilt adaptivity
 * (astg_adjust_e) keeps them. */
#define ASTG_FIELD_DENSITY 0
#define ASTG_FIELD_HEAT 1
int astg_set_scalar(astg_ctx* ctx, int field, int level, const double* oct, const double* tilt);
int astg_get_scalar(astg_ctx* ctx, int field, int level, double* oct, double* tilt);

/* The chain between project and advect in OctreeOmniGridSolver3::onUpdate
 * (src/octree_omnigrid_solver3.cpp:234-241): constrainVelocity on the MAC and
 * tilt grids (closed walls; no collider), velocity->updateGhost, updateDualGrid,
 * dual constraint, dual->updateGhost. */
int astg_post_project(astg_ctx* ctx);
/* OctreeOmniGrid3<REAL>::updateGhost of one scalar field (octree_omni_grid3.h:136-143). */
int astg_scalar_update_ghost(astg_ctx* ctx, int field);
/* Dual velocity grid of the last astg_post_project: 3 values per node for the
 * cell centres and for the vertices. */
int astg_get_dual(astg_ctx* ctx, int level, double* oct3, double* tilt3);

/* Point queries (host arrays, n points of 3 coordinates):
 *   ASTG_SAMPLE_VELOCITY  OctreeOmniFaceCenteredGrid3::sample (omni_face_centered_grid3.h:37-80)
 *   ASTG_SAMPLE_MAC       OctreeFaceCenteredGrid3::sample (face_centered_grid3.cpp:36-45)
 *   ASTG_SAMPLE_BACKTRACE BackTrace over dt (advection.h:15-64)
 *   ASTG_SAMPLE_DENSITY / ASTG_SAMPLE_HEAT  OctreeOmniGrid3<REAL>::sample (omni_grid3.h:146-191)
 *   ASTG_SAMPLE_SDF       boundary SDF, CellCenteredGrid::sample (cell_centered_grid.h:40-54)
 * out holds 3 values per point for the first three, 1 for the others. */
#define ASTG_SAMPLE_VELOCITY 0
#define ASTG_SAMPLE_MAC 1
#define ASTG_SAMPLE_BACKTRACE 2
#define ASTG_SAMPLE_DENSITY 3
#define ASTG_SAMPLE_HEAT 4
#define ASTG_SAMPLE_SDF 5
int astg_sample_points(astg_ctx* ctx, int what, int n, const double* xyz, double dt, double* out);

/* astg_sample_points without any allocation: the same queries, results and
 * errors, with the points and results staged through a fixed device array
 * (one per GPU, 12 MiB, held until the process ends) instead of buffers the
 * context allocates on demand. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (tight serving loops, harnesses
 * that forbid hipMalloc). At most 262144 points per call (ASTG_ERR_ARG
 * beyond); the context's fields must already be on the device (ASTG_ERR_STATE
 * otherwise). All calls on one GPU share the array: calls that overlap in
 * time on one GPU, from any context or thread, race. */
int astg_sample_points_noalloc(astg_ctx* ctx, int what, int n, const double* xyz, double dt, double* out);

/* Scene source of the reference's octree smoke example (sphereSource,
 * example/octree_smoke_3d/main.cpp:23-81) with its constants as parameters:
 * inside the sphere (centre, radius) density and heat are set on cell and
 * vertex DOFs (and half-tilts), MAC z-faces get `speed`, every leaf tilt whose
 * centre is inside gets speed/sqrt(3) on all 8 slots, half-tilts their 4 tilt
 * slots (and their face slot for z-faces). update_ghost: then update the
 * density and heat ghosts (the example does so on its first call). The
 * example's solver calls it at the start of every step (sceneSetup). */
int astg_apply_sphere_source(astg_ctx* ctx, double cx, double cy, double cz, double radius, double density,
                             double heat, double speed, int update_ghost);

/* OctreeOmniGridSolver3::wipeBoundary (src/octree_omnigrid_solver3.cpp:1240-1292):
 * density and heat at leaf cell centres, leaf vertices and half-tilts where
 * the boundary SDF is negative become 0 (onUpdate runs it after sceneSetup,
 * before addForce). A no-op without a collider. */
int astg_wipe_boundary(astg_ctx* ctx);

/* The example's sphere source with an inflow vector instead of the rising
 * speed: inside the sphere density and heat as astg_apply_sphere_source, every
 * MAC face gets its component of velocity[3], every tilt slot of a leaf tilt
 * inside and the tilt and face slots of a half-tilt the projection of
 * velocity on the slot's direction. No counterpart in the reference, whose
 * source rises along +z only (example/octree_smoke_3d/main.cpp:23-81): it
 * drives the two lateral sources of the full-scale configuration. */
int astg_apply_sphere_source_velocity(astg_ctx* ctx, double cx, double cy, double cz, double radius, double density,
                                      double heat, const double* velocity, int update_ghost);

/* OctreeOmniGridSolver3::addForce (src/octree_omnigrid_solver3.cpp:917-972):
 * buoyancy f = smoke_density_factor * density + temperature_factor * (heat -
 * ambT), ambT the mean cell heat (nodes of levels above 0 weighted 8), sampled
 * at every valid MAC z-face, open tilt slot and half-tilt slot and added along
 * +z over dt. The reference's defaults are -0.000625 and 5.0
 * (octree_omnigrid_solver3.h:125-126). */
int astg_add_force(astg_ctx* ctx, double dt, double smoke_density_factor, double temperature_factor);

/* OctreeOmniGridSolver3::adjustE (src/octree_omnigrid_solver3.cpp:1125-1166):
 * every non-fixed leaf tilt closes when open and its density leaves
 * (d_threshold, max_threshold), or opens when closed, the velocity at the tilt
 * exceeds v_threshold and the density sampled at pos - v*dt lies inside the
 * range; then the new tilt sizes (h/2 or 0) and ghost states are committed.
 * The numbering, weights and projection pattern follow the new state; fields
 * are kept and sampled with the previous tilt state until the next
 * astg_advect, as the reference's tiltE/tTiltE pair does. */
int astg_adjust_e(astg_ctx* ctx, double d_threshold, double dt, double v_threshold, double max_threshold);

/* enableTiltAdaptivity (include/octree_omnigrid_solver3.h:56), off by
 * default: when on, astg_advect and astg_step start the advection with
 * adjustE(0.02, dt, DBL_EPSILON, 1000) (octree_omnigrid_solver3.cpp:852-853). */
int astg_set_tilt_adaptivity(astg_ctx* ctx, int on);

/* OctreeOmniGridSolver3::advect (src/octree_omnigrid_solver3.cpp:850-914):
 * with tilt adaptivity on, adjustE first; then density and heat on octagon,
 * open tilt and half-tilt DOFs, their ghosts, the velocity on MAC, tilt and
 * half-tilt slots (SemiLagrangianOctreeOmniFace3, advection.h:195-239), then
 * the 0.999 decay of both scalars. Needs astg_post_project first. */
int astg_advect(astg_ctx* ctx, double dt);

/* One step of onUpdate after sceneSetup and addForce (which the caller runs
 * first: astg_apply_sphere_source, astg_add_force): constrain + ghosts,
 * astg_project, astg_post_project, astg_advect. */
int astg_step(astg_ctx* ctx, double dt, double rho, const astg_pcg_params* params, astg_pcg_result* result);

/* ---- differentiable step ------------------------------------------------ */
/* One step of onUpdate as a differentiable map on a frozen topology without a
 * collider (no source, no force, no tilt adaptivity): constrain + ghosts,
 * project, post-projection chain, advect (density and velocity; the heat is
 * advected too but does not enter the state). Its state is the flat vector
 * MAC 3G | tilt 8G | dual cell 3G | dual vertex 3G | density cell G |
 * density vertex G (G and node order as astg_field_layout); entries a step
 * does not write carry over. No counterpart in the reference (the reference
 * has no adjoint). */
int astg_diff_state_size(astg_ctx* ctx, long long* n);
/* State of the context to / from caller memory (device_pointers as in
 * astg_fields, in `stream` order). */
int astg_diff_get_state(astg_ctx* ctx, double* state, int device_pointers, void* stream);
int astg_diff_set_state(astg_ctx* ctx, const double* state, int device_pointers, void* stream);
/* The step's forward on the context's state. */
int astg_diff_step(astg_ctx* ctx, double dt, double rho, const astg_pcg_params* params, astg_pcg_result* result);
/* Adjoint of one step: adj_in = J^T adj_out, J the step's Jacobian at
 * state_in. The transposed solve reuses the MIC(0)-PCG (A is symmetric);
 * the fixed linear passes are applied transposed from records made once per
 * topology; the advection's transpose scatters through its interpolation
 * weights and back-propagates through the back-traced positions. Leaves the
 * context's state at state_in after its projection. */
int astg_diff_step_adjoint(astg_ctx* ctx, double dt, double rho, const astg_pcg_params* params,
                           const double* state_in, const double* adj_out, double* adj_in, int device_pointers,
                           void* stream);
/* One operator of the step, forward (transpose = 0: y = op(x)) or transposed
 * (y = op^T x), on host state vectors; for checks. ASTG_OP_ADVECT is
 * linearised at the context's current state (its transpose is J^T there). */
#define ASTG_OP_CONSTRAIN 0    /* wall constraint and velocity ghosts */
#define ASTG_OP_POST 1         /* post-projection chain (constraint, ghosts, dual grid, dual ghosts) */
#define ASTG_OP_SCALAR_GHOST 2 /* density ghosts */
#define ASTG_OP_PROJECT 3      /* projection (assembly of b, solve, velocity correction) */
#define ASTG_OP_ADVECT 4       /* advection of density and velocity, density ghosts, decay */
int astg_diff_apply(astg_ctx* ctx, int op, int transpose, double dt, double rho, const astg_pcg_params* params,
                    const double* x, double* y);

/* ---- frames on disk (the reference's binary formats) ------------------- */
/* Writes the context's state as the reference's frame dumps
 * (src/viewer/grid_dump.cpp): layout (dumpOctreeLayout3dBinary, 280-327),
 * tilt state (dumpOctreeOmniTiltE3dBinary, 448-467), density and heat
 * (dumpOctreeOmniScalar3dBinary, 379-410), velocity
 * (dumpOctreeOmniFaceCentered3dBinary, 492-549). A null path skips that file;
 * layout and tilt need only a topology, the fields a device context. */
int astg_write_frame(astg_ctx* ctx, const char* layout, const char* tilt, const char* density, const char* heat,
                     const char* velocity);
/* restart (src/octree_omnigrid_solver3.cpp:1015-1057) on a context whose
 * topology was set from the same split map: the layout file must describe
 * the context's layers exactly (ASTG_ERR_IO otherwise; the reference reads it
 * into the layers it already built); the tilt state is replaced as restart
 * replaces tiltE, keeping the numbering's own open / fixed state as restart
 * keeps tilt_layout; density, heat and velocity are loaded, then velocity
 * ghosts, the dual grid and its ghosts are recomputed (no constraint), as
 * restart does. Null paths skip; every file is checked before state changes. */
int astg_read_frame(astg_ctx* ctx, const char* layout, const char* tilt, const char* density, const char* heat,
                    const char* velocity);

/* ---- uniform-layout AST projection ------------------------------------ */
/* The reference's uniform-grid AST solver, ast::ASTSolver<d> (d = 2 or 3,
 * include/ast_solver.h:20-673; the pybind module ast_solver exposes it as
 * ASTSolver2 / ASTSolver3, python_bindings/ast_solver.cc:43-54), projection
 * path: ASTSolver::project (ast_solver.h:322-498) with the reference's build
 * configuration (CORRECTION_MAC, no edge-node unknowns, ast_general.h:11-15).
 * Unknowns: every cell in box order (x fastest), then every open tilt node in
 * box order (updateIndex, 599-621). Grids in the reference's layouts, x
 * fastest: tilt nodes prod(res + 1), MAC faces of axis a prod(res + e_a)
 * (FaceCenteredGrid::vecIndex, face_centered_grid.h:72-80), tilt velocities
 * node-major with 4 (2D) or 8 (3D) slots per node. Host pointers. */
typedef struct astg_ast astg_ast;

/* ASTSolver::initialize's solver settings (ast_solver.h:40): tol 1e-5 as an
 * ABSOLUTE bound on ||r||_inf (the reference's std::vector PCGSolver::solve,
 * pcg_solver.h:316, does not scale it by ||b||), 5000 iterations, MIC(0) 0.97,
 * min-diagonal ratio 0.95, FP64, reference numbering. */
void astg_ast_params_default(astg_pcg_params* p);

/* ASTSolver(res, dx) + initialize() (ast_solver.h:27-51): fields zero, every
 * tilt closed. res[d] >= 2 each; in 3D res[1] <= res[0] (constrainVelocity
 * runs x over [0, res_y) on the y walls, ast_solver.h:634, and writes past its
 * face grid otherwise). On failure *out is still set: read astg_ast_last_error,
 * then destroy it. */
int astg_ast_create(astg_ast** out, int device, int dim, const int* res, double dx);
int astg_ast_destroy(astg_ast* ast);
const char* astg_ast_last_error(const astg_ast* ast);
/* The tilt layout (TiltNode::open per node, ast_grid_layout.h:13-60): one
 * flag per tilt node; boundary nodes are held closed (initData 79-87,
 * constrainFixedTiltNode). */
int astg_ast_set_layout(astg_ast* ast, const uint8_t* tilt_open);
/* ASTFaceGrid's MAC and tilt data (ast_face_grid.h:16-58); mac_z unused in 2D. */
int astg_ast_set_velocity(astg_ast* ast, const double* mac_x, const double* mac_y, const double* mac_z,
                          const double* tilt);
int astg_ast_get_velocity(astg_ast* ast, double* mac_x, double* mac_y, double* mac_z, double* tilt);
/* project(dt) (ast_solver.h:322-498): numbering, assembly, MIC(0)-PCG,
 * correctVelocityByPressure, constrainVelocity; rho is the solver's rho
 * (2.0 in the reference). params may be null (astg_ast_params_default).
 * ASTG_ORDER_BLOCK_COLOUR (ASTG_ORDER_AUTO with the mixed precision) solves in
 * block colours: squares / cubes of 4 cells in 4 / 8 colours, an added form
 * with another preconditioner; ASTG_TAU_AUTO is 0.97 in FP64, 0 (IC(0)) in
 * the mixed form. */
int astg_ast_project(astg_ast* ast, double dt, double rho, const astg_pcg_params* params, astg_pcg_result* result);
/* The last projection's system (CSR, rows sorted by column), right-hand side
 * and pressure; n and nnz as in its result. Any pointer may be null. */
int astg_ast_system(astg_ast* ast, int n, long long nnz, int* rowstart, int* col, double* val, double* b, double* p);
/* The unknown of every tilt node (pIndex, -1 where closed). */
int astg_ast_tilt_index(astg_ast* ast, int* tilt_index);

#ifdef __cplusplus
}
#endif

#endif /* ASTG_H */
