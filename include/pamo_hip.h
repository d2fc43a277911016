/* C ABI of the MI355X (gfx950) stage-2 simplifier, libpamo_hip.so.
 *
 * Every pointer named d_* is a device pointer on the context's GPU; every
 * stream argument is a hipStream_t passed as void* (NULL = default stream).
 * Functions return 0 on success and a negative code on error; the message is
 * available from pamo_last_error().  No function aborts the process.
 *
 * Ownership: inputs are borrowed and copied into context buffers, so the
 * caller's buffers are never modified (as simp_cuda/src/cusimp_free.cu:926-946
 * copies them).  Output views point into context memory and stay valid until
 * the next call on the same context.  A context is not thread-safe; two
 * contexts may be used from two threads.
 */
#ifndef PAMO_HIP_H
#define PAMO_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pamo_ctx pamo_ctx;

/* shuffle_mode for the init round: 1 = thrust's Feistel shuffle with
 * minstd_rand seed 1 (what simp_cuda/src/cusimp_free.cu:937-941 calls),
 * 0 = identity. */
#define PAMO_SHUFFLE_IDENTITY 0
#define PAMO_SHUFFLE_FEISTEL 1

/* Replaces the CUDSP_Free constructor (simp_cuda/src/pybind.cpp:16-20,191-193):
 * creates the context that owns all device state of one simplifier. */
int pamo_ctx_create(pamo_ctx** out, int device, int shuffle_mode);

/* Replaces ~CUDSP_Free (simp_cuda/src/pybind.cpp:21-49): syncs and frees. */
void pamo_ctx_destroy(pamo_ctx* ctx);

/* Last error message of the context (empty string when none). */
const char* pamo_last_error(const pamo_ctx* ctx);

/* Device bytes the context holds: its buffers, the batched engine's and the
 * results not yet taken, and the stage-3 system's (the fixed region's size
 * for a context in the no-allocation mode).  No counterpart in the
 * reference; for callers that admit work by HBM footprint.  -1 for a null
 * context. */
long long pamo_device_bytes(pamo_ctx* ctx);

/* Replaces CUDSP_Free.forward / CUSimp_Free::forward
 * (simp_cuda/src/pybind.cpp:52-72, simp_cuda/src/cusimp_free.cu:891-1048):
 * one guarded simplification round.  d_points [n_pts,3] f32, d_tris
 * [n_tris,3] i32 (rows of -1 are removed faces), d_undo [n_undo] i32 (the
 * vertex undo list returned by the previous round). */
int pamo_free_forward(pamo_ctx* ctx, void* stream, const float* d_points, const int* d_tris, const int* d_undo,
                      int n_undo, int n_pts, int n_tris, float scale, float threshold, int is_stuck, int init);

/* Replaces CUDSP.forward / CUSimp::forward (simp_cuda/src/pybind.cpp:140-181,
 * simp_cuda/src/cusimp.cu:587-720), the unguarded round that PaSP runs:
 * fans, edges, quadrics, costs (no invalid-vertex table, the fan length term
 * counted twice, no x5 skinny weight), claims and collapses without the
 * skinny restore, then the vertex map.  No self-intersection check, undo
 * list, invalid-face removal or line-edge collapse.  Outputs as for
 * pamo_free_forward (points, tris, occupancy, map; the undo list is empty).
 * Use one context per simplifier object: the edge costs persist between
 * calls, as in the reference. */
int pamo_sp_forward(pamo_ctx* ctx, void* stream, const float* d_points, const int* d_tris, int n_pts, int n_tris,
                    float scale, float threshold, int init);

/* The five outputs of the last forward (simp_cuda/src/pybind.cpp:74-102):
 * points [n_pts,3], tris [n_tris,3], occupancy [n_pts], vertex map [n_pts],
 * vertex undo list [n_undo].  Device views into the context. */
int pamo_free_outputs(pamo_ctx* ctx, float** d_points, int** d_tris, int** d_occ, int** d_map, int** d_undo,
                      int* n_pts, int* n_tris, int* n_undo);

/* Copies the five outputs of the last forward into caller device buffers
 * (the clones of simp_cuda/src/pybind.cpp:74-100); any pointer may be NULL. */
int pamo_free_copy_outputs(pamo_ctx* ctx, void* stream, float* d_points, int* d_tris, int* d_occ, int* d_map,
                           int* d_undo);

/* The stage-2 loop of PaMO.run (simp_cuda/pamo/__init__.py:121-158) kept on
 * the device: rounds of forward + compaction until the face count reaches
 * target (or 10), or stays unchanged `tolerance` times.  d_verts are the
 * centred vertices, scale the largest axis extent.  On return out_V and out_F
 * hold the final sizes; fetch the mesh with pamo_run_fetch.
 * A negative target disables the size exits (target and <= 10 faces), as in
 * the Blender add-on's loop.
 * stats (may be NULL): [0] rounds [1] exit reason (1 target, 2 <=10 faces,
 * 3 tolerance, 4 iteration cap) [2] collapses [3] undo passes (max)
 * [4] self-intersection checks [5] query faces [6] candidate pairs
 * [7] intersecting pairs found (engine; the round-by-round loop counts
 * intersecting query faces). */
int pamo_run_stage2(pamo_ctx* ctx, void* stream, const float* d_verts, const int* d_faces, int n_verts, int n_faces,
                    float scale, long long target, float threshold, int tolerance, long long iters, int* out_V,
                    int* out_F, long long* stats);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc): the loop of
 * pamo_run_stage2, with the result copied into caller buffers d_verts_out
 * [n_verts,3] and d_faces_out [n_faces,3] (sizes as pamo_run_stage2 reports
 * them in out_V / out_F), and every working buffer carved from one fixed
 * device region of PAMO_NOALLOC_SCRATCH_BYTES that the library holds on each
 * GPU it is loaded on until the process ends.  The call allocates no device
 * or pinned host memory.  Meshes above PAMO_NOALLOC_MAX_VERTS vertices or
 * PAMO_NOALLOC_MAX_FACES faces are rejected (-4, message in
 * pamo_last_error), as is a context with the strict guard set.  ctx supplies
 * the device, the shuffle mode and the error message; its own buffers are
 * not used.  All calls on one GPU share its region: calls from threads of
 * one process are serialised by a lock, and the call synchronises the
 * stream before it returns; work of the caller's that still reads or writes
 * the region (there is none unless the caller takes its address) would race
 * with the next call. */
#define PAMO_NOALLOC_SCRATCH_BYTES (128u << 20)
#define PAMO_NOALLOC_MAX_FACES 65536
#define PAMO_NOALLOC_MAX_VERTS 65536
int pamo_run_stage2_noalloc(pamo_ctx* ctx, void* stream, const float* d_verts, const int* d_faces, int n_verts,
                            int n_faces, float scale, long long target, float threshold, int tolerance,
                            long long iters, float* d_verts_out, int* d_faces_out, int* out_V, int* out_F,
                            long long* stats);

/* The same loop as pamo_run_stage2 driven round by round from the host over
 * pamo_free_forward (the exact call sequence of simp_cuda/pamo/__init__.py:
 * 131-158: forward, then the compaction of lines 140-148).  Same arguments,
 * outputs and stats; kept as the second implementation the device loop is
 * tested against. */
int pamo_run_stage2_rounds(pamo_ctx* ctx, void* stream, const float* d_verts, const int* d_faces, int n_verts,
                           int n_faces, float scale, long long target, float threshold, int tolerance,
                           long long iters, int* out_V, int* out_F, long long* stats);

/* Face count after each round of the last pamo_run_stage2 (host array of
 * cap entries); returns the number of rounds. */
int pamo_run_rounds(pamo_ctx* ctx, int* host_faces, int cap);

/* Copies the mesh of the last pamo_run_stage2 into caller device buffers. */
int pamo_run_fetch(pamo_ctx* ctx, void* stream, float* d_verts_out, int* d_faces_out);

/* Replaces pamo_simplify of the Blender add-on (pamo_blender/
 * cusimpinterface.cu:57-161, bound with ctypes in pamo_blender/__init__.py:
 * 75-79): host arrays in, host arrays out, one process-wide context on the
 * current device.  in_out_nv_ni = {vertices, faces} on entry and the result
 * sizes on return; out_verts/out_tris must hold the input sizes.  Runs the
 * PaMO.run loop on the raw (uncentred) vertices with scale = the largest
 * bounding-box extent, tolerance 4, no target and no size exit.  As in the
 * reference, errors are printed to stderr; the sizes are then set to -1. */
void pamo_simplify(float* in_verts, int* in_tris, float* out_verts, int* out_tris, int* in_out_nv_ni, float threshold);

/* The centring and scale of PaMO.run (simp_cuda/pamo/__init__.py:60-64,117,
 * 126) on host arrays, with numpy's float32 arithmetic: tris_mean = the
 * sequential float32 sum from +0 over faces of (((0 + a) + b) + c) / 3,
 * divided by F (the order numpy's mean(axis=1).mean(axis=0) takes),
 * verts_out = points -
 * tris_mean, *scale = the largest extent of verts_out.  Host only (no GPU
 * call); face indices must lie in [0, V).  Returns 0, or -1 for F <= 0 or
 * V <= 0. */
int pamo_prepare_host(const float* points, const int* faces, int V, int F, float* verts_out, float* tris_mean,
                      float* scale);

/* Batched stage 2 (no counterpart in the reference, which simplifies one mesh
 * per PaMO.run call, simp_cuda/pamo/__init__.py:107-165).  Many meshes share
 * one packed device state; every round of pamo_batch_step runs the loop body
 * of PaMO.run for all active meshes in shared launches, and each mesh's
 * result is bit-identical to running it alone with pamo_run_stage2.
 *
 * pamo_batch_submit copies one mesh in (centred vertices, faces, scale and
 * the run parameters of pamo_run_stage2; job is an opaque caller tag) and
 * returns its handle (>= 0).  The mesh joins at the next round, also while
 * others are mid-run.  The copies are stream-ordered (as hipMemcpyAsync):
 * d_verts and d_faces must stay unchanged until later work on the stream,
 * or a synchronisation, has passed the call.  Face indices are checked
 * before the call returns (-2 when one lies outside [0, n_verts)). */
int pamo_batch_submit(pamo_ctx* ctx, void* stream, const float* d_verts, const int* d_faces, int n_verts, int n_faces,
                      float scale, long long target, float threshold, int tolerance, long long iters, long long job);

/* Runs up to max_rounds rounds of every active mesh; meshes that meet their
 * exit condition are retired between rounds.  Returns the number retired;
 * n_active receives the meshes still running. */
int pamo_batch_step(pamo_ctx* ctx, void* stream, int max_rounds, int* n_active);

/* Finished, not yet taken jobs: writes up to cap (handle, job) entries and
 * returns how many there are. */
int pamo_batch_poll(pamo_ctx* ctx, int* handles, long long* jobs, int cap);

/* Result sizes of a job (returns 1 while it still runs).  stats: [0] rounds
 * [1] exit reason [2] collapses [3] undo passes (max) [4] job tag [5] input
 * vertices [6] input faces [7] microseconds from admission to retirement;
 * round_faces receives up to cap per-round face counts. */
int pamo_batch_result(pamo_ctx* ctx, int handle, int* n_verts, int* n_faces, long long* stats, int* round_faces,
                      int cap);

/* Copies a finished job's mesh into caller device buffers and frees it. */
int pamo_batch_take(pamo_ctx* ctx, void* stream, int handle, float* d_verts_out, int* d_faces_out);

/* Engine totals: [0] active meshes [1] packed vertices [2] packed faces
 * [3] self-intersection checks [4] query faces [5] candidate pairs. */
int pamo_batch_state(pamo_ctx* ctx, long long out[6]);

/* Kernel timing (no counterpart in the reference, whose only timing is
 * wall-clock prints, simp_cuda/pamo/__init__.py:121,160-163): HIP events
 * around the selected kernels on the call's stream.  pamo_profile resets the
 * totals and selects kernels by mask (1 << PAMO_K_*); pamo_profile_read
 * returns launches, total milliseconds and total algorithmic bytes. */
#define PAMO_K_QUADRIC 0
#define PAMO_K_EDGE_COST 1
#define PAMO_K_CLAIM 2
#define PAMO_K_LBVH 3
#define PAMO_K_TRAVERSE 4
int pamo_profile(pamo_ctx* ctx, unsigned kernel_mask);
int pamo_profile_read(pamo_ctx* ctx, int kernel, double out[3]);

/* Parity hooks: copies an internal buffer of the last forward to host memory
 * (returns the byte size; copies only when cap_bytes is large enough).
 * which: 5 fan offsets, 6 fan, 7 edges, 8 quadrics (16 f32 per vertex),
 * 9 edge costs, 10 per-face claim keys, 11 collapsed-edge list,
 * 12 first check's intersecting-face flags, 13 round-start faces,
 * 14 invalid-vertex table. */
long long pamo_debug_copy(pamo_ctx* ctx, void* stream, int which, void* host_dst, long long cap_bytes);

/* Test hook (no counterpart in the reference): limits the candidate-pair
 * buffer of the engine's self-intersection check to cap pairs (-1 restores
 * the default of 12 per face).  Queries whose pairs do not fit are walked by
 * the single-kernel fallback; results must not change. */
int pamo_debug_pair_cap(pamo_ctx* ctx, long long cap);

/* Test hook (no counterpart in the reference): how the engine's
 * self-intersection check finds candidate pairs.  0 (default) by mesh size:
 * every query against every alive face of its mesh when all meshes have at
 * most 32768 faces, else an LBVH walk; 1 always the LBVH; 2 always the
 * exhaustive search.  Both give the same candidate set. */
int pamo_debug_guard(pamo_ctx* ctx, int mode);

/* Test hook (no counterpart in the reference): the engine's LBVH walk splits
 * each query over 2^kshift lanes, one subtree each (by size: 3 up to 65,536
 * alive faces, 2 up to 262,144, else 0).  kshift 0..3 forces a split, -1
 * restores the choice by size.  The candidate set does not depend on it. */
int pamo_debug_walk_split(pamo_ctx* ctx, int kshift);

/* Strict guard for the engine (pamo_run_stage2, pamo_batch_*); no
 * counterpart in the reference, off by default.  For callers that need every
 * output to pass an exact self-intersection check: a candidate pair also
 * counts as intersecting when the exact closed triangle test (filtered FP64
 * orientation predicates with an exact fallback) says so, which catches the
 * pairs the reference's narrow phase skips (coplanar, shared-edge
 * coordinates, short shared-vertex contacts); the undo scan covers every
 * collapsed edge, and undo passes repeat until no collapse is marked (at
 * most 64).  Results then differ from the reference's wherever it keeps an
 * intersection.  The round kernels (pamo_free_forward) are unaffected. */
int pamo_set_strict(pamo_ctx* ctx, int strict);

/* Launch shapes of the engine, by name (no counterpart in the reference,
 * whose BLOCK_SIZE 512 and 10-bit Morton resolution are compiled in,
 * simp_cuda/src/cusimp_free.cu:14, bvh/morton_code.cuh:20).  None changes a
 * result.  Names: "ec_lanes" (lanes per edge of the cost kernel: 2, 4, 8),
 * "morton_bits" (per axis, 6..10), "bf_max_faces" (meshes up to this many
 * faces use the exhaustive search, larger ones the LBVH),
 * "bf_split_max_faces", "walk8_max", "walk4_max", "emit_groups", "pipeline"
 * (1: the next round is queued before this one's result is read; 0: never),
 * "graph" and "graph_max_faces" (rounds of one mesh of at most that many
 * faces are recorded into hipGraphs and replayed; 0: never), and for
 * stage 3 "s3_bf_max_faces" (meshes up to this many faces search contacts
 * exhaustively, larger ones through LBVHs), "s3_cg_one_block" (meshes up
 * to this many vertices run each CG step in one workgroup) and "s3_graph"
 * (each Newton iteration after its contact detection is recorded into a
 * hipGraph and replayed; 0: never).  A context
 * starts from the defaults, then from the flat JSON file named by the
 * environment variable PAMO_TUNING_FILE, if set (python -m pamo_amd.tune
 * writes one).  Returns 0, or -1 for an unknown name or a value out of
 * range. */
int pamo_set_tuning(pamo_ctx* ctx, const char* name, long long value);
int pamo_get_tuning(pamo_ctx* ctx, const char* name, long long* value);

/* Placement for the engine (pamo_run_stage2, pamo_batch_*,
 * pamo_run_stage2_noalloc); no counterpart in the reference, which places
 * and scores every collapse at the edge midpoint (simp_cuda/src/
 * cusimp_free.cu:491,687-688).  0 (default): the midpoint.  1: the optimal
 * vertex of Q_u + Q_v, for callers that want lower geometric error:
 * the 3x3 system solved in FP32 when well conditioned, else in FP64, else
 * (or when the solution lies farther from the midpoint than the edge is
 * long) the one of midpoint, u and v with the lowest quadric error.  The
 * cost is scored at that position and the collapse moves u there.  The
 * round kernels (pamo_free_forward) are unaffected. */
int pamo_set_placement(pamo_ctx* ctx, int mode);

/* Stage 1 (remeshing), the two calls of simp_cuda/pamo/__init__.py:75-93.
 * pamo_mesh_to_sdf replaces torchcumesh2sdf.get_sdf(tris, R, band)
 * (__init__.py:82): d_tris holds n_tris triangles as [n_tris, 3, 3] floats
 * in [0, 1]^3; d_volume receives the R^3 signed distance lattice (sample
 * (i, j, k) at ((i + 0.5) / R, ...), index (k R + j) R + i): exact distance
 * to the nearest triangle within band, +-band beyond it, positive outside
 * (samples connected to the lattice boundary through samples at least
 * 0.75 / R from the surface), negative inside.
 * pamo_volume_to_mesh replaces pdmc.DMC(volume) (__init__.py:48,85): the
 * zero level set of an R^3 volume (positive: outside) as a closed,
 * consistently oriented (outward) 2-manifold, vertices in the lattice's
 * [0, 1]^3 coordinates; the counts are returned and the mesh is kept until
 * pamo_volume_mesh_fetch copies it to caller buffers (n_verts x 3 floats,
 * n_faces x 3 ints).  R is at most PAMO_REMESH_MAX_R.  Returns 0, -4 for
 * invalid arguments, -1 for a device error. */
#define PAMO_REMESH_MAX_R 512
int pamo_mesh_to_sdf(pamo_ctx* ctx, void* stream, const float* d_tris, int n_tris, int R, float band,
                     float* d_volume);
int pamo_volume_to_mesh(pamo_ctx* ctx, void* stream, const float* d_volume, int R, int* n_verts, int* n_faces);
int pamo_volume_mesh_fetch(pamo_ctx* ctx, void* stream, float* d_verts, int* d_faces);

/* Stage 3 (safe projection): the Stage3System of simp_cuda/safe_project/src/
 * pamo_safe_project/system.py:17-470 on the device, in float64, one system
 * per context.  All arrays are device pointers.
 * pamo_s3_configure replaces Stage3Config (config.py:6-73): cfg_d = {d_hat,
 * coll_stiffness, contact_detection_radius, ee_classify_thres,
 * ccd_slackness, ccd_thickness, elas_young_modulus, elas_poisson_ratio,
 * hinge_stiffness, mesh2gt_dist_stiffness, min_angle_thres}, cfg_i =
 * {n_newton_iters, n_cg_iters, n_line_search_iters, spd_max_iters,
 * ccd_max_iters, max_blocks}; either may be NULL for the defaults.
 * pamo_s3_register replaces register_mesh(gt_V, gt_F, V, F) (system.py:
 * 210-320): the target mesh (n_gt_verts x 3 doubles, n_gt_faces x 3 ints)
 * and the mesh to project (closed, consistently oriented: one hinge per
 * edge, energy.py:626-628; -2 otherwise).  pamo_s3_step replaces n calls of
 * step() (system.py:407-436).  pamo_s3_get / pamo_s3_set copy an
 * n_verts x 3 vector: 0 positions (get_vertices, system.py:438-440),
 * 1 targets, 2 forces (-dE/dx), 3 Hessian diagonal, 4 Newton step,
 * 5 rest positions, 6 positions at the Newton iteration's start, 7 at the
 * last contact detection.  pamo_s3_sizes: n_verts, n_faces, n_edges,
 * n_hinges, n_contacts, do_nothing.  pamo_s3_scalars (host out[5]):
 * energy, energy before the line search, CCD step, detection radius,
 * halvings taken.  Returns 0, -4 for invalid arguments, -1 for an error. */
int pamo_s3_configure(pamo_ctx* ctx, const double* cfg_d, const long long* cfg_i);
int pamo_s3_register(pamo_ctx* ctx, void* stream, const double* d_gt_verts, int n_gt_verts, const int* d_gt_faces,
                     int n_gt_faces, const double* d_verts, int n_verts, const int* d_faces, int n_faces);
int pamo_s3_step(pamo_ctx* ctx, void* stream, int n_steps);
int pamo_s3_sizes(pamo_ctx* ctx, long long* out);
int pamo_s3_get(pamo_ctx* ctx, void* stream, int which, double* d_out);
int pamo_s3_set(pamo_ctx* ctx, void* stream, int which, const double* d_in);
int pamo_s3_scalars(pamo_ctx* ctx, void* stream, double* out);
/* Stage-3 parity hooks (the parts of a step, system.py:407-436):
 * pamo_s3_update_target (_update_vertex_target), pamo_s3_newton_part
 * (0 contact detection, 1 energy, 2 derivatives, 3 CG solve, 4 step clamp,
 * 5 CCD, 6 line search, 7 one whole Newton iteration), the energy's four
 * parts at positions d_x (host parts[4]: target distance, membrane, hinge,
 * contact), the contact list as rows (type, i0, i1, i2, i3, distance type)
 * with distances (returns the count; writes when d_rows is not NULL), H dx,
 * and the topology: edges (n_edges x 2), hinges (x0, x1, x2, x3 per edge),
 * rest angle and length per hinge, Voronoi area per vertex. */
int pamo_s3_update_target(pamo_ctx* ctx, void* stream);
int pamo_s3_newton_part(pamo_ctx* ctx, void* stream, int part);
int pamo_s3_energy_parts(pamo_ctx* ctx, void* stream, const double* d_x, double* parts);
long long pamo_s3_contacts(pamo_ctx* ctx, void* stream, int* d_rows, double* d_dist);
int pamo_s3_hess_dx(pamo_ctx* ctx, void* stream, const double* d_dx, double* d_out);
int pamo_s3_topology(pamo_ctx* ctx, void* stream, int* d_edges, int* d_hinges, double* d_rest, double* d_vor);

/* Parity hook: every ordered candidate pair (query, current) of one
 * self-intersection check over all alive faces of a mesh, with the
 * reference narrow-phase verdict (simp_cuda/src/bvh/self_intersect.cuh:151-392).
 * host_pairs receives rows (query, current, hit); returns the pair count
 * (writes only when cap_rows is large enough), or a negative error. */
long long pamo_debug_pairs(pamo_ctx* ctx, void* stream, const float* d_points, const int* d_tris, int n_pts,
                           int n_tris, int* host_pairs, long long cap_rows);

#ifdef __cplusplus
}
#endif
#endif
