/*
 * fos.h -- C ABI of libfos.so, the MI355X (gfx950) implementation of the
 * per-frame foam step of VirtualOrganics/Fabric-of-Space-Taichi-JFA.
 *
 * Every entry point takes raw device pointers (contiguous, caller-owned),
 * sizes, scalars and a hipStream_t passed as void* (NULL = default stream).
 * Work is enqueued on that stream; functions that return values on the host
 * (pointers documented "host") synchronise that stream first.  Nothing is
 * allocated per call: scratch lives in a context made by fos_ctx_create.
 *
 * Return value: 0 on success, FOS_EINVAL (-22) for bad arguments, or
 * -(1000 + hipError_t) for a HIP runtime error.  fos_last_error() gives the
 * text of the last failure on the calling thread.
 *
 * Layouts: pos float[n][3] (x, y, z), rad float[n], labels int32[res^3]
 * indexed x*res*res + y*res + z, per-particle rows int32/float[n][32],
 * hash cells int32[G^3] indexed cx*G*G + cy*G + cz.
 *
 * Each declaration cites the reference function (file:line) it replaces.
 */
#ifndef FOS_H
#define FOS_H
#include <stdint.h>

#define FOS_OK 0
#define FOS_EINVAL (-22)
#define FOS_ABI_VERSION 2

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fos_ctx fos_ctx;
typedef struct fos_batch fos_batch;

int fos_abi_version(void);
const char *fos_last_error(void);

/* Launch tunables (process-wide):
 *   "jfa_variant"      -1..3  JFA pass kernel: bit 0 evaluates four candidate
 *                      list entries per loop iteration (two otherwise), bit 1
 *                      carries each voxel's winning power distance from pass
 *                      to pass (needs a context with max_res <= 256); -1
 *                      (default) picks per context: four entries when ids
 *                      exceed 65,535, the carry whenever available
 *   "face_emit_tx"     0,1,2,4,8,16  face-pair emission: 0 merges pairs
 *                      within a wave; others merge within a workgroup tile of that
 *                      many x-planes in LDS (default 8)
 *   "face_emit_zw"     16,32,64  z-extent of the emission tile (its y-extent
 *                      is 256 / zw rows; default 32)
 *   "grid_block"       0,1  hash-grid rebuild: 1 (default) one workgroup
 *                      per range of cells, independent of each other; 0 the
 *                      count/scan/scatter/rank kernels
 *   "grid_range_cells" 0..1024  cells per workgroup of the range rebuild of
 *                      one foam (0, default: C/192 clamped to [64, 1024])
 *   "pbd_lanes"        0,8,16,32  lanes per particle of the slot-record PBD
 *                      pass (fos_pbd_pass_cells with records,
 *                      fos_batch_pbd_pass); 0 (default): 32 below 131,072
 *                      particles, else 8
 *   "pressure_flow"    0,1  1 (default): fos_pressure_run runs its iterations
 *                      in one launch where it can; 0: one launch per
 *                      iteration
 *   "fin_first_burst"  0,2..16  finalize rounds queued before the first host
 *                      read of the fixed-point flags (0, default: as many as
 *                      the context's last finalize needed)
 * They change speed only, never results.  No reference counterpart (Taichi
 * picks launch shapes itself). */
int fos_set_param(const char *name, int value);
int fos_get_param(const char *name, int *value);

/* Context: scratch for up to max_n particles, res <= max_res voxels per axis,
 * max_cells hash cells.  Replaces the field allocation of jfa.init_jfa
 * (jfa.py:120-204), grid/dynamics fields (run.py:129-199). */
int fos_ctx_create(fos_ctx **out, int max_n, int max_res, int max_cells);
int fos_ctx_destroy(fos_ctx *ctx);

/* Device fills (Taichi field clears, e.g. grid.py:146-153 clear_grid). */
int fos_fill_i32(int32_t *dst, int64_t count, int32_t value, void *stream);
int fos_fill_f32(float *dst, int64_t count, float value, void *stream);

/* ---- jfa.py ------------------------------------------------------------ */

/* fos_face_accumulate then fos_finalize_faces (jfa.py:703-857) with one host
 * round trip in the usual case: emission, compaction and a burst of
 * fixed-point rounds are queued together and one read sees both the
 * row-table overflow and the fixed-point flags (an overflow regrows the
 * tables and starts over before anything the caller keeps is written). */
int fos_jfa_faces(fos_ctx *ctx, int res, int n, const int32_t *filtered, const float *rad,
                  float voxel, int32_t *face_ids, int32_t *face_counts, float *prev_scores,
                  int32_t *neighbor_list, int32_t *neighbor_count, int32_t *used_count,
                  int32_t *fsc, void *stream);
/* The tail of a JFA frame in fewer launches: fos_jfa_faces, then
 * fos_fsc_ema (jfa.py:766-785, inside the finalize's last launch), then
 * fos_fsc_gate + fos_fsc_targets_gated (run.py:1184-1231) as one launch;
 * *gate (device float) receives the warm-start gate (1.0 = growth).  Results
 * equal the four calls bit for bit.  No reference counterpart (launches
 * saved per frame). */
int fos_jfa_faces_controller(fos_ctx *ctx, int res, int n, const int32_t *filtered,
                             float *rad_target, const float *rad, float voxel, int32_t *face_ids,
                             int32_t *face_counts, float *prev_scores, int32_t *neighbor_list,
                             int32_t *neighbor_count, int32_t *used_count, int32_t *fsc,
                             float *fsc_ema, float alpha, int frame, int warmstart_frames,
                             double fsc_zero_rate, int f_low, int f_high, float growth_pct,
                             float r_min, float r_max, float deadband, float *gate,
                             void *stream);
/* Per-frame seed table (x, y, z, (beta*r)^2) that the JFA kernels gather;
 * the reference recomputes the same terms per candidate (jfa.py:470-477). */
int fos_jfa_seed_table(fos_ctx *ctx, const float *pos, const float *rad,
                       int n, float beta, void *stream);
/* jfa.py:485-520 jfa_init: clears grid and temp (res^3) to -1, builds the
 * seed table, seeds every particle's voxel (argmin power distance, lowest
 * index on ties).  With n > 0 it also clears the context's pass counters, so
 * the next fos_jfa_passes on the same stream skips that clear. */
int fos_jfa_init(fos_ctx *ctx, const float *pos, const float *rad, int n,
                 int res, float L, float voxel, float beta, int32_t *grid,
                 int32_t *temp, void *stream);
/* jfa.py:524-623 jfa_step_ping0/1: one pass in -> out at step_size over the
 * seed table built by the last fos_jfa_init/fos_jfa_seed_table; adds the
 * changed-voxel count into *changes (device int32). */
int fos_jfa_step(fos_ctx *ctx, int res, float L, float voxel, int step_size,
                 const int32_t *in, int32_t *out, int32_t *changes,
                 void *stream);
/* jfa.py:626-632 jfa_copy_temp_to_grid. */
int fos_jfa_copy(int res, const int32_t *src, int32_t *dst, void *stream);
/* jfa.py:1078-1116 run_jfa pass loop with the early exit decided on the
 * device (no host sync).  info (device int32[2 + num_passes]) receives
 * [0] passes executed, [1] early-exit flag, [2+p] changes of pass p.  The
 * final labels end in grid. */
int fos_jfa_passes(fos_ctx *ctx, int res, float L, float voxel,
                   int num_passes, int early_exit, double th_pass_delta,
                   int min_passes, int32_t *grid, int32_t *temp,
                   int32_t *info, void *stream);
/* fos_jfa_passes that also records events[p] (hipEvent_t) on the stream
 * before pass p and events[num_passes] after the last pass, so each pass
 * kernel can be timed on the stream it runs on; NULL entries record nothing
 * (events[0] and events[num_passes] alone time the whole schedule without a
 * record between the passes).  No reference counterpart:
 * per-pass timing for the benchmark (the reference times JFA on the host,
 * run.py:1071-1080). */
int fos_jfa_passes_timed(fos_ctx *ctx, int res, float L, float voxel,
                         int num_passes, int early_exit, double th_pass_delta,
                         int min_passes, int32_t *grid, int32_t *temp,
                         int32_t *info, void *const *events, void *stream);
/* Limits of fos_jfa_labels_noalloc's static scratch. */
#define FOS_STATIC_MAX_N 262144
#define FOS_STATIC_MAX_RES 384
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under graph capture, serving loops, harnesses that forbid
 * hipMalloc).  Computes what fos_jfa_init (grid only, temp untouched before
 * the passes) followed by fos_jfa_passes computes -- the seeding and the
 * whole pass schedule, the final labels in grid and info as fos_jfa_passes
 * writes it -- with no context: the seed table and the pass counters live
 * in one static device array (about 4.5 MB per device, held until the
 * process ends).  Calls no hipMalloc/hipFree.  Rejects
 * n > FOS_STATIC_MAX_N or res > FOS_STATIC_MAX_RES with FOS_EINVAL.  Every
 * call on one device shares that device's array: calls that overlap on one
 * device race. */
int fos_jfa_labels_noalloc(const float *pos, const float *rad, int n, int res, float L,
                           float voxel, float beta, int num_passes, int early_exit,
                           double th_pass_delta, int min_passes, int32_t *grid, int32_t *temp,
                           int32_t *info, void *stream);
/* The static scratch limits (host): particles, resolution, bytes. */
int fos_static_limits(int *max_n, int *max_res, int64_t *bytes);
/* hipEvent_t helpers for fos_jfa_passes_timed: create count events, destroy
 * them, and read the count-1 elapsed intervals (ms) between consecutive
 * events after waiting for the last one. */
int fos_events_create(int count, void **events);
int fos_events_destroy(int count, void **events);
int fos_events_elapsed(int count, void *const *events, float *ms);
/* jfa.py:639-699 majority_filter_labels: grid -> filtered. */
int fos_majority(int res, const int32_t *grid, int32_t *filtered,
                 void *stream);
/* jfa.py:702-762 accumulate_face_area: face_ids/face_counts rows [n][32] in
 * first-encounter order (serial semantics).  Synchronises the stream once to
 * check the row tables for overflow. */
int fos_face_accumulate(fos_ctx *ctx, int res, int n, const int32_t *filtered,
                        int32_t *face_ids, int32_t *face_counts,
                        void *stream);
/* jfa.py:788-857 finalize_faces_kernel (+ add_neighbor jfa.py:400-443) with
 * the serial-order hysteresis resolved by a fixed point over rows.
 * prev_scores persists across frames.  Synchronises per fixed-point round. */
int fos_finalize_faces(fos_ctx *ctx, int n, const float *rad, float voxel,
                       const int32_t *face_ids, const int32_t *face_counts,
                       float *prev_scores, int32_t *neighbor_list,
                       int32_t *neighbor_count, int32_t *used_count,
                       int32_t *fsc, void *stream);
/* jfa.py:765-785 update_fsc_ema. */
int fos_fsc_ema(int n, float alpha, const int32_t *fsc, float *fsc_ema,
                void *stream);
/* jfa.py:1204-1330 compute_pair_symmetry / compute_overflow_count /
 * compute_fsc_stats / compute_score_stats.  Host out_i[5] = {p_dir,
 * overflow, min_fsc, max_fsc, sum_fsc}; host out_f[4] = {mean_fsc,
 * mean_fsc_ema, mean_score, std_score}. */
int fos_jfa_stats(fos_ctx *ctx, int n, const float *rad, float voxel,
                  const int32_t *fsc, const float *fsc_ema,
                  const int32_t *used_count, const int32_t *face_ids,
                  const int32_t *face_counts, int64_t *out_i, double *out_f,
                  void *stream);
/* jfa.py:1337-1397 check_symmetry / check_overflow / check_self_loops.
 * Host out[3] = {asymmetric, overflow, self_loops}. */
int fos_validate(fos_ctx *ctx, int n, const int32_t *neighbor_list,
                 const int32_t *neighbor_count, int64_t *out, void *stream);
/* jfa.py:966-1008 mark_dirty_tiles (tile_dirty uint8[tpa^3]). */
int fos_dirty_mark(const float *pos, const float *rad, const float *pos_prev,
                   const float *rad_prev, int n, int res, float L,
                   float voxel, int tile, int halo, float pos_thr,
                   float rad_thr, uint8_t *tile_dirty, void *stream);
/* jfa.py:1011-1023 count_dirty_tiles (host *count). */
int fos_dirty_count(fos_ctx *ctx, const uint8_t *tile_dirty, int tiles,
                    int64_t *count, void *stream);
/* jfa.py:1036-1049 update_tile_cache. */
int fos_tile_cache_update(const float *pos, const float *rad, int n,
                          float *pos_prev, float *rad_prev, void *stream);

/* ---- grid.py ----------------------------------------------------------- */

/* grid.py:189-222 count_particles_per_cell (cell_count must be cleared). */
int fos_grid_count(const float *pos, int n, int G, float L,
                   int32_t *cell_count, void *stream);
/* grid.py:229-243 prefix_sum: exclusive scan over G^3 cells. */
int fos_grid_scan(fos_ctx *ctx, const int32_t *cell_count,
                  int32_t *cell_start, int G, void *stream);
/* grid.py:268-309 scatter_particles: stores wrapped positions back into pos
 * and writes cell_indices, ascending particle index within each cell
 * (stable).  cell_write holds cell_start on entry and the cell ends on exit,
 * as in the reference. */
int fos_grid_scatter(float *pos, int n, int G, float L, int32_t *cell_write,
                     int32_t *cell_indices, void *stream);
/* run.py:476-504 rebuild_grid: clear, count, scan, copy, scatter. */
int fos_grid_build(fos_ctx *ctx, float *pos, int n, int G, float L,
                   int32_t *cell_count, int32_t *cell_start,
                   int32_t *cell_indices, void *stream);

/* ---- dynamics.py ------------------------------------------------------- */

/* fos_fsc_targets applied only when *gate (device) > 0.5 (the gate that
 * fos_fsc_gate computed on the device; no host round trip). */
int fos_fsc_targets_gated(const float *rad, float *rad_target, const int32_t *fsc,
                          const float *fsc_ema, int n, int f_low, int f_high, float growth_pct,
                          float r_min, float r_max, float deadband, const float *gate,
                          void *stream);
/* run.py:1184-1195 on the device: *gate = 0 while frame < warmstart_frames
 * or the fraction of particles with FSC 0 exceeds fsc_zero_rate, else 1. */
int fos_fsc_gate(fos_ctx *ctx, const int32_t *fsc, int n, int frame, int warmstart_frames,
                 double fsc_zero_rate, float *gate, void *stream);
/* dynamics.py:77-171 project_overlaps, Jacobi form: reads pos_in, writes
 * pos_out (distinct buffers).  gap_factor = (float)(1 + GAP_FRACTION). */
int fos_pbd_pass(const float *pos_in, float *pos_out, const float *rad,
                 int n, int G, float L, float gap_factor, float max_disp_frac,
                 float disp_mult, const int32_t *cell_start,
                 const int32_t *cell_count, const int32_t *cell_indices,
                 void *stream);
/* fos_pbd_pass followed by the first step of the fos_grid_build that comes
 * next in the frame (grid.py:285-309 position wrap and cell index, kept in
 * the context), fused into the PBD kernel; fos_grid_build_from_cells then
 * completes that rebuild.  cell_pr (may be NULL): the slot records of the
 * current CSR (fos_grid_build_pr / fos_grid_build_from_cells), read instead
 * of gathering each neighbour's position and radius by index; particles are
 * then taken in cell order.  Results equal fos_pbd_pass + fos_grid_build
 * bit for bit.  No reference counterpart (a launch saved per PBD pass). */
int fos_pbd_pass_cells(fos_ctx *ctx, const float *pos_in, float *pos_out, const float *rad,
                       int n, int G, float L, float gap_factor, float max_disp_frac,
                       float disp_mult, const int32_t *cell_start, const int32_t *cell_count,
                       const int32_t *cell_indices, const float *cell_pr, void *stream);
/* The rest of fos_grid_build (grid.py:189-309 count, prefix sum, scatter)
 * from the cells fos_pbd_pass_cells left in the context.  cell_pr (may be
 * NULL, else pos and rad are read): cell_pr[k] = (position, radius) of the
 * particle in slot k, float4 per slot. */
int fos_grid_build_from_cells(fos_ctx *ctx, const float *pos, const float *rad, int n, int G,
                              int32_t *cell_count, int32_t *cell_start, int32_t *cell_indices,
                              float *cell_pr, void *stream);
/* fos_grid_build (grid.py:30-309) that also writes the slot records
 * cell_pr[k] = (wrapped position, radius) of the particle in slot k (cell_pr
 * may be NULL).  No reference counterpart. */
int fos_grid_build_pr(fos_ctx *ctx, float *pos, const float *rad, int n, int G, float L,
                      int32_t *cell_count, int32_t *cell_start, int32_t *cell_indices,
                      float *cell_pr, void *stream);
/* dynamics.py:304-372 compute_local_max_overlaps + reduce_max_depth: fills
 * local_max_depth and the device float *max_depth (a true maximum). */
int fos_max_overlap(const float *pos, const float *rad, int n, int G, float L,
                    const int32_t *cell_start, const int32_t *cell_count,
                    const int32_t *cell_indices, float *local_max_depth,
                    float *max_depth, void *stream);
/* dynamics.py:1242-1320 set_fsc_targets. */
int fos_fsc_targets(const float *rad, float *rad_target, const int32_t *fsc,
                    const float *fsc_ema, int n, int f_low, int f_high,
                    float growth_pct, float r_min, float r_max,
                    float deadband, void *stream);
/* dynamics.py:1322-1363 compute_controller_stats.  Host out_i[3] = {grow,
 * shrink, idle}; host out_f[2] = {mean_dr, max_abs_dr}. */
int fos_controller_stats(fos_ctx *ctx, const float *rad,
                         const float *rad_target, const float *fsc_ema, int n,
                         int f_low, int f_high, int64_t *out_i, double *out_f,
                         void *stream);
/* dynamics.py:1366-1427 nudge_radii_adaptive_ema. */
int fos_nudge_radii(float *rad, const float *rad_target,
                    const float *local_max_depth, int n, float alpha,
                    float max_step_pct, float r_min, float r_max,
                    float back_global, int mode, float contact_tol,
                    void *stream);
/* run.py:733-751 settle step with its counter on the device, so the frame
 * needs no host read before it: the counter entering the frame is
 * adjustment_frames when force (the last frame set it without a gate), else
 * *settle_in, reset to adjustment_frames when apply_gate and the last
 * frame's device gate maxd[1] (run.py:1184-1231, fos_fsc_gate) is > 0.5.  A
 * positive counter runs nudge_radii_adaptive_ema (dynamics.py:1366-1427;
 * mode 1 takes the global backpressure run.py:744-749 from the overlap
 * maximum maxd[0], computed in double with contact_tol_d) and counts down.
 * The counter leaving the frame is written to *settle_out (!= settle_in). */
int fos_settle_nudge(float *rad, const float *rad_target, const float *local_max_depth, int n,
                     float alpha, float max_step_pct, float r_min, float r_max, int mode,
                     float contact_tol, double contact_tol_d, const float *maxd,
                     const int32_t *settle_in, int32_t *settle_out, int adjustment_frames,
                     int force, int apply_gate, void *stream);
/* dynamics.py:1454-1557 equilibrate_pressure, one Jacobi iteration: delta_r
 * from the current radii, then the clamped apply.  stats (device float[2])
 * receives {max|dr|, changed count as float bits of an int32} and is
 * overwritten by this call. */
int fos_pressure_iter(int n, int frame, float *rad, float *delta_r,
                      const int32_t *face_ids, const int32_t *fsc, float k,
                      float p_exp, float pair_cap, int max_nei, float r_min,
                      float r_max, float *stats, void *stream);
/* `iters` calls of equilibrate_pressure (dynamics.py:1454-1557) at one frame,
 * each a Jacobi iteration; the result ends in rad, delta_r holds the last
 * iteration's deltas and stats (device float[2], optional) its
 * {max|dr|, changed count as int32 bits}.  The per-frame loop of run.py:755-771
 * repeated `iters` times.  With 2 <= iters <= 32, max_nei <= 16,
 * n <= min(max_n, 131072), a grid the device holds at once and the stream
 * not capturing, all iterations run in one launch in which each particle
 * waits only on its own partners' previous iteration (16 lanes per particle,
 * or one for foams too large for that; tunable "pressure_flow"; the stream
 * must run nothing concurrently with it); otherwise one kernel per iteration
 * over a ping-pong buffer.  Both give
 * the same bits.  Calls on one context must not overlap. */
int fos_pressure_run(fos_ctx *ctx, int n, int frame, int iters, float *rad,
                     float *delta_r, const int32_t *face_ids, const int32_t *fsc,
                     float k, float p_exp, float pair_cap, int max_nei, float r_min,
                     float r_max, float *stats, void *stream);
/* dynamics.py:1560-1589 compute_pressure_stats: host out[2] = {rmin, rmax}. */
int fos_pressure_stats(fos_ctx *ctx, const float *rad, int n, float *out,
                       void *stream);
/* No counterpart in the reference: a self-check of the fast path that the
 * pressure kernels take for x^(1/P) at P = 3 (the float the double pow gives,
 * decided from a Newton cube root when its rounding is unambiguous).  For the
 * positive floats with bit patterns first + k * stride, k < count (all below
 * 0x80000000), adds to the device counters counts[0] the arguments where the
 * fast path returns another float than (float)pow((double)x, fp32(1/3)) and
 * to counts[1] those it leaves to the double pow.  The caller zeroes counts. */
int fos_pow_third_check(unsigned first, unsigned count, unsigned stride,
                        unsigned long long *counts, void *stream);
/* dynamics.py:1596-1634 apply_brownian_motion with Philox4x32-10 noise keyed
 * (seed, frame, particle).  mode 0: uniform (u-0.5)*2*strength (reference
 * form); mode 1: Gaussian with sigma = strength. */
int fos_brownian(float *pos, int n, float strength, float L, uint64_t seed,
                 int frame, int mode, void *stream);
/* fos_brownian followed by the first step of the fos_grid_build that comes
 * next in the frame (grid.py:285-309 position wrap and cell index into a
 * G^3 grid, kept in the context); fos_grid_build_from_cells completes that
 * rebuild.  Results equal fos_brownian + fos_grid_build bit for bit.  No
 * reference counterpart (a launch saved per frame). */
int fos_brownian_cells(fos_ctx *ctx, float *pos, int n, float strength, float L, uint64_t seed,
                       int frame, int mode, int G, void *stream);
/* dynamics.py:607-660 apply_brownian (Ornstein-Uhlenbeck velocity update with
 * Philox uniforms keyed (seed, frame, particle)) fused with integrate_jitter;
 * pos == NULL updates v_jit only.  mean_r is the host-computed mean radius
 * (run.py:860-861); drift_cap_frac = (float)(MAX_DRIFT_FRACTION*GAP_FRACTION). */
int fos_ou_jitter(float *pos, float *v_jit, const float *rad, int n,
                  float mean_r, float dt, float jitter_rms, float jitter_tau,
                  float drift_cap_frac, float L, uint64_t seed, int frame,
                  void *stream);
/* fos_ou_jitter with the mean radius computed on the device (run.py:860-861
 * computes it on the host): the exact double sum of the radii over n, rounded
 * to fp32, sets the noise scale; no host read.  mean_out (device float, may be
 * NULL) receives that mean. */
int fos_ou_jitter_mean(fos_ctx *ctx, float *pos, float *v_jit, const float *rad, int n,
                       float dt, float jitter_rms, float jitter_tau, float drift_cap_frac,
                       float L, uint64_t seed, int frame, float *mean_out, void *stream);
/* dynamics.py:649-660 integrate_jitter: pos = wrapP(pos + v_jit*dt). */
int fos_integrate_jitter(float *pos, const float *v_jit, int n, float dt,
                         float L, void *stream);
/* dynamics.py:532-593 apply_xsph_smoothing. */
int fos_xsph(const float *pos, float *vel, float *vel_tmp, int n, int G,
             float L, float lim_sq, float eps, const int32_t *cell_start,
             const int32_t *cell_count, const int32_t *cell_indices,
             void *stream);
/* dynamics.py:473-480 apply_global_damping. */
int fos_damp(float *vel, int n, float factor, void *stream);

/* ---- one whole frame ----------------------------------------------------- */

/* The frame of run.py:700-1261 (plain JFA path) as one call: what
 * foam.FoamSim.step() issues, in the same order, with the same host reads.
 * fos_frame_desc holds the configuration, fos_frame_bufs the caller-owned
 * device arrays (pos[2] is a ping-pong pair: state.pos_cur says which holds
 * the positions; maxd_host is pinned host memory of 2 floats), and
 * fos_frame_state what carries from frame to frame (the caller zeroes it
 * once, sets need_rebuild when it wrote the positions since the last step,
 * and calls fos_frame_release when done).  After a step: pbd_passes,
 * max_depth and jfa_ran describe it; gate_read says whether the previous
 * gated frame's warm-start gate was read (its value in gate); pending says
 * whether this frame's gate is still on the device (maxd[1]).  pass_events
 * (may be NULL) as for fos_jfa_passes_timed.  No reference counterpart (the
 * reference's loop is Python). */
typedef struct fos_frame_desc {
  int n, res, G, jfa_passes;
  float L, voxel, beta, ema_alpha;
  int early_exit, min_passes;
  double th_pass_delta;
  int pbd_base, pbd_max;
  double pbd_scale, r_max_d;
  float gap_factor, max_disp_frac, disp_mult;
  float alpha, max_step_pct, r_min, r_max, contact_tol;
  double contact_tol_d;
  int backpressure_mode, adjustment_frames, warmstart_frames;
  double fsc_zero_rate;
  int f_low, f_high;
  float growth_pct, deadband;
  int pressure, pressure_iters, pressure_max_nei;
  float pressure_k, pressure_exp, pressure_cap;
  int brownian, brownian_mode;
  float brownian_strength;
  uint64_t seed;
  int ou_jitter;
  float dt, jitter_rms, jitter_tau, drift_cap_frac;
  int jfa_every, jfa_warm, slot_records;
  /* which frames run the JFA: 0 the cadence above (jfa_every, jfa_warm),
   * 1 the reference's temporal skip (fos_jfa_schedule, run.py:902-1018) */
  int schedule;
  int skip_warmstart, max_skip;       /* JFA_WARMSTART_FRAMES, MAX_SKIP_FRAMES */
  double th_rsigma, th_margin, th_fsc_drift, oob_frac;  /* TH_*, out-of-band guard */
  /* adaptive resolution on JFA frames (run.py:1049-1068): res =
   * round(L_d * voxels_per_diameter / (2 r_mean)) clamped to [res_min,
   * res_max] (res_max <= the context's max_res) */
  int adaptive_res, res_min, res_max;
  double voxels_per_diameter, L_d;
} fos_frame_desc;
typedef struct fos_frame_bufs {
  float *pos[2];
  float *rad, *rad_target, *lmd, *delta_r, *v_jit;
  int32_t *cell_count, *cell_start, *cell_indices;
  float *cell_pr;
  int32_t *grid, *temp, *filtered, *face_ids, *face_counts, *fsc;
  float *fsc_ema;
  int32_t *neighbor_list, *neighbor_count, *used_count;
  float *prev_scores;
  int32_t *info;
  float *maxd, *maxd_host;
  int32_t *settle_dev;
  /* scheduler (schedule 1 or adaptive_res): positions and radii at the last
   * decision, and pinned host memory of 8 doubles for the statistics */
  float *pos_prev, *rad_prev;
  double *sched_host;
} fos_frame_bufs;
typedef struct fos_frame_state {
  int frame, pos_cur, sidx, settle_force, settle_gate, need_rebuild, pending;
  int pbd_passes, jfa_ran, gate_read;
  float max_depth, gate;
  void *event;
  /* scheduler state (run.py:165-172, 904-1018): decisions made, frames since
   * the last JFA, whether pos_prev/rad_prev hold a decision's values,
   * whether last_fsc_mu is set (and whether the next decision sets it: the
   * FSC mean after a JFA), the band-change lockout, the last decision's
   * reason (FOS_JFA_* below) and signals (sigma(|dr|)/r_mean, max |dpos|,
   * |d mean FSC|, out-of-band share, r_mean), and the resolution the last
   * JFA frame ran at (0: desc->res) */
  int sched_frames, since, have_prev, fsc_mu_set, fsc_mu_pending, lockout, reason, jfa_res;
  double last_fsc_mu;
  double signals[5];
} fos_frame_state;
/* reasons of a scheduling decision, in the reference's priority order */
#define FOS_JFA_SKIP 0
#define FOS_JFA_WARMSTART 1
#define FOS_JFA_BAND_CHANGE 2
#define FOS_JFA_WATCHDOG 3
#define FOS_JFA_OUT_OF_BAND 4
#define FOS_JFA_ACTIVITY 5
#define FOS_JFA_CADENCE 6
int fos_frame_step(fos_ctx *ctx, const fos_frame_desc *d, const fos_frame_bufs *b,
                   fos_frame_state *s, void *const *pass_events, void *stream);
/* `frames` calls of fos_frame_step in a loop, frame s->frame onward, with
 * pass_events[k] (pass_events or its entries may be NULL) for frame k of the
 * run.  The same calls in the same order, so the same bits; the host issues
 * the next frame while the GPU still runs the end of this one.  State after
 * the run as after its last step.  No reference counterpart. */
int fos_frame_run(fos_ctx *ctx, const fos_frame_desc *d, const fos_frame_bufs *b,
                  fos_frame_state *s, int frames, void *const *const *pass_events,
                  void *stream);
/* Frees what fos_frame_step created in the state (its host-read event). */
int fos_frame_release(fos_frame_state *s);
/* run.py:902-1018 signal measurement of the temporal-skip scheduler, on the
 * device, with rad_prev/pos_prev then set to rad/pos (the reference's
 * rad_prev_cached and pos_prev): out[0] mean |rad - rad_prev|, out[1] its
 * population standard deviation, out[2] max |pos - pos_prev| at the minimum
 * image (|d| > L/2 -> d - sign(d) L; fp32 as numpy computes it), out[3]
 * mean radius, out[4] mean FSC, out[5] particles with fsc < f_low or fsc >
 * f_high.  Sums in fp64 in a fixed order (deterministic).  With have_prev 0
 * out[0..2] are 0.  out: device memory of 8 doubles. */
int fos_sched_stats(fos_ctx *ctx, const float *pos, const float *rad, const int32_t *fsc, int n,
                    float L, int f_low, int f_high, float *pos_prev, float *rad_prev,
                    int have_prev, double *out, void *stream);
/* The JFA decision of frame s->frame, made where the reference makes it
 * (after the dynamics, run.py:902-1018): with desc->schedule 1 the
 * temporal-skip triggers (warm start, watchdog, band-change lockout,
 * activity, out-of-band guard) from fos_sched_stats and one host read, else
 * the cadence; with desc->adaptive_res the resolution of a JFA frame from
 * the mean radius (run.py:1049-1068).  Sets *run, *res and the state's
 * scheduler fields; pos is the current positions.  No reference counterpart
 * as an entry point (the reference decides in its Python loop). */
int fos_jfa_schedule(fos_ctx *ctx, const fos_frame_desc *d, const fos_frame_bufs *b,
                     fos_frame_state *s, const float *pos, int *run, int *res, void *stream);

/* ---- topology.py ------------------------------------------------------- */

/* topology.py:68-133 build_topo_candidates: pairs[n][cap], count[n],
 * truncated[n], in the reference's cell/occupant order. */
int fos_topo_candidates(const float *pos, const float *rad, int n, int G,
                        float L, float mult, int cap,
                        const int32_t *cell_start, const int32_t *cell_count,
                        const int32_t *cell_indices, int32_t *pairs,
                        int32_t *count, int32_t *truncated, void *stream);

/* topology.py:140-262 gabriel_test_topological_degree: clears deg, then for
 * every stored candidate pair (pairs[n][cap], pair_count) that passes the
 * hash subsampling (stride > 1) and the high-degree early exit, adds 1 to
 * both degrees when no other particle lies strictly inside the diameter
 * sphere.  cell_size is the configured CELL_SIZE (stencil sizing). */
int fos_topo_gabriel(const float *pos, int n, int G, float L, float cell_size,
                     const int32_t *cell_start, const int32_t *cell_count,
                     const int32_t *cell_indices, const int32_t *pairs, const int32_t *pair_count,
                     int cap, int frame, int stride, int early_exit, float deg_high,
                     const float *deg_ema, int32_t *deg, void *stream);
/* topology.py:335-446 gabriel_update_batched: particles i % num_batches ==
 * batch_idx rebuild their candidates (j > i, 1/q of pairs by hash) and add
 * their Gabriel pairs to deg (not cleared).  pair_count counts past cap as
 * the reference's counter does; only stored pairs are tested. */
int fos_topo_gabriel_batched(const float *pos, const float *rad, int n, int G, float L,
                             float cell_size, float mult, int cap, const int32_t *cell_start,
                             const int32_t *cell_count, const int32_t *cell_indices,
                             int batch_idx, int num_batches, int q, int frame, int32_t *pairs,
                             int32_t *pair_count, int32_t *truncated, int32_t *deg,
                             void *stream);
/* topology.py:458-504 knn_topo_degree: deg[i] = min(k, neighbours within
 * (mult*ri)^2 over the 27 cells). */
int fos_topo_knn(const float *pos, const float *rad, int n, int G, float L, float mult, int k,
                 const int32_t *cell_start, const int32_t *cell_count,
                 const int32_t *cell_indices, int32_t *deg, void *stream);
/* topology.py:266-283 update_topo_ema and :449-455 commit_topo_to_ema:
 * ema = alpha*deg + (1-alpha)*ema. */
int fos_topo_ema(const int32_t *deg, float *ema, int n, float alpha, void *stream);
/* topology.py:287-322 update_radii_topological. */
int fos_topo_radii(float *rad, const float *deg_eff, int n, float deg_low, float deg_high,
                   float gain_grow, float gain_shrink, float r_min, float r_max, void *stream);

/* ---- learned face predictor -------------------------------------------
 * No reference counterpart: the reference always runs the JFA
 * (jfa.run_jfa, jfa.py:1056-1197).  These score candidate pairs for a shared
 * power-diagram face on frames where the JFA is decimated; the plain JFA
 * stays the judge (see the frame driver). */

/* Candidate pairs: every j != i within mult*(ri+rj) over the 27 hash cells
 * around i's cell (cell size >= mult*2*r_max; the CSR of fos_grid_build),
 * in cell/occupant order (the shape of topology.build_topo_candidates,
 * topology.py:68-133), the first `cap` (<= 128) kept.  Row i of
 * cand[n][cap] lists the kept partners j > i first (nup[i] of them, scan
 * order), then those j < i (scan order), -1 padded; ncand[n] counts both,
 * truncated[n] = 1 when more were found.  When rel is given, rel[n][cap][4]
 * holds beside each kept partner k (pk - pi at the minimum image,
 * |pk - pi|^2 - rk^2). */
int fos_pred_candidates(const float *pos, const float *rad, int n, int G, float L, float mult,
                        int cap, const int32_t *cell_start, const int32_t *cell_count,
                        const int32_t *cell_indices, int32_t *cand, int32_t *ncand, int32_t *nup,
                        float *rel, int32_t *truncated, void *stream);
/* Flat list of the scored pairs (i, cand[i][s]), s < nup[i] (each
 * unordered pair once, in the row of its lower index): pair_off[n+1] is the
 * exclusive prefix of nup (pair_off[n] = number of pairs, on the device),
 * pair_i[p] the particle of pair p (capacity sum of nup <= n*cap). */
int fos_pred_pairs(fos_ctx *ctx, const int32_t *nup, int n, int32_t *pair_off, int32_t *pair_i,
                   void *stream);
/* 13 features per pair p < min(pair_off[n], max_pairs) into
 * feat[max_pairs][16] (entries 13-15 zero) from the rel rows of
 * fos_pred_candidates; r_mean normalises the radii, clip_r*(ri+rj) is the
 * half-size of the square clipped to the face polygon; feature 12 is 1 when
 * the partner is in row i of prev_up[n][32] (the previous frame's accepted
 * partners j > i; may be NULL: 0). */
int fos_pred_features(const float *rad, int n, float r_mean, float clip_r, const int32_t *cand,
                      const int32_t *ncand, const float *rel, const int32_t *pair_off,
                      const int32_t *pair_i, int cap, int64_t max_pairs, const int32_t *prev_up,
                      float *feat, void *stream);
/* prev_up[n][32] from a JFA frame: row i = the partners j > i of i's
 * neighbour list (jfa.py:789-857 output), in list order, -1 padded. */
int fos_pred_prev_rows(const int32_t *neighbor_list, const int32_t *neighbor_count, int n,
                       int32_t *prev_up, void *stream);
/* MLP 12->64->64->1 on MFMA with BF16 weights/activations and FP32
 * accumulation: logit[p] for p < npairs, or p < min(npairs, *npairs_dev)
 * when npairs_dev (device) is given.  w1_packed / w2_packed hold the BF16
 * weights in MFMA lane order (face_predictor.pack). */
int fos_pred_mlp(const float *feat, int64_t npairs, const void *w1_packed, const void *w2_packed,
                 const float *b1, const float *b2, const float *w3, float b3, float *logit,
                 const int32_t *npairs_dev, void *stream);
/* Pair p = (i, cand[i][p - pair_off[i]]) is kept when logit[p] > 0 and
 * counts in fsc of both particles (fsc[n] is overwritten); optional face
 * rows (kept partners j > i in candidate order, at most 32, -1 padded;
 * counts 1) for the pressure step (jfa.py:702-762 row shape). */
int fos_pred_fsc(const float *logit, const int32_t *cand, const int32_t *pair_off, int n, int cap,
                 int32_t *fsc, int32_t *face_ids, int32_t *face_counts, void *stream);
/* Training label of each pair p < min(pair_off[n], max_pairs): 1 when its
 * partner is in i's JFA neighbour list, else 0. */
int fos_pred_labels(const int32_t *cand, const int32_t *pair_off, const int32_t *pair_i, int n,
                    int cap, int64_t max_pairs, const int32_t *neighbor_list,
                    const int32_t *neighbor_count, float *labels, void *stream);
/* Number of FP32 MLP parameters (flat layout W1[16][64] b1 W2[64][64] b2 w3
 * b3) into host *count. */
int fos_mlp_param_count(int *count);
/* One Adam step on a batch of B sample indices with class-weighted binary
 * cross entropy.  Forward, backward and weight-gradient GEMMs run on MFMA
 * with BF16 operands (X, W1, W2, H1, dZ2, dZ1) and FP32 accumulation; bias
 * and last-layer gradients, the loss and Adam are FP32.  scratch holds at
 * least FOS_MLP_TRAIN_SCRATCH floats (per-wave partial gradients); *loss
 * (device) is increased by the batch mean loss.  Bit-reproducible for a
 * given B. */
#define FOS_MLP_TRAIN_SCRATCH (64 * (5313 + 1))
int fos_mlp_train_step(float *params, float *grad, float *adam_m, float *adam_v, int step,
                       float lr, const float *feat, const float *labels, const int32_t *batch_idx,
                       int B, float pos_weight, float *scratch, float *loss, void *stream);

/* ---- batches of foams ---------------------------------------------------
 * No reference counterpart (the reference runs one foam per process).  A
 * batch concatenates B independent foams of uneven size: foam b owns
 * particles [p0_b, p0_b + n_b) of the batch arrays (p0 = prefix of n), labels
 * [v0_b, v0_b + res_b^3) (v0 = prefix of res^3) and hash cells
 * [c0_b, c0_b + G_b^3).  Particle and label ids are batch-global.  Each phase
 * of the frame runs as one launch over all foams.  A context used with a
 * batch must hold its total particle count. */

/* Builds the descriptors and work tables of B foams: particle counts n,
 * JFA grid edges res (voxel L/res), scheduled passes (<= 15), hash grid
 * edges G, radius bounds, FSC bands and noise seeds. */
int fos_batch_create(fos_batch **out, int B, const int *n, const int *res, const int *passes,
                     const int *G, const float *r_min, const float *r_max, const float *f_low,
                     const float *f_high, const unsigned *seed, float L);
int fos_batch_destroy(fos_batch *batch);
/* Host out[7] = {B, total particles, total voxels, total cells, max res,
 * max passes, max G}. */
int fos_batch_info(const fos_batch *batch, int64_t *out);
/* JFA pass tile of each foam (host int[B]): 0 rows (a workgroup = 4 z-rows
 * of 64 voxels; lanes past res idle), 1 flat (256 consecutive voxels of a
 * y-z plane; row offsets per lane).  Results do not depend on it; the
 * ensemble autotuner times both per grid edge.  Synchronises (tables are
 * copied to the device). */
int fos_batch_set_tiles(fos_batch *batch, const int *kinds);
/* jfa.py:485-520 for every foam: labels -1, seed table, seeding (lowest
 * global index on ties = lowest index within the foam). */
int fos_batch_jfa_init(fos_batch *batch, fos_ctx *ctx, const float *pos, const float *rad,
                       float L, float beta, int32_t *grid, int32_t *temp, void *stream);
/* jfa.py:1082-1116 for every foam, one launch per pass over all foams; a
 * foam stops at its own schedule or early exit.  info[B][2 + 16] (device):
 * passes executed, exit flag, changed voxels per pass; labels end in grid. */
int fos_batch_jfa_passes(fos_batch *batch, fos_ctx *ctx, float L, int early_exit,
                         double th_pass_delta, int min_passes, int32_t *grid, int32_t *temp,
                         int32_t *info, void *stream);
/* jfa.py:640-699 for every foam. */
int fos_batch_majority(fos_batch *batch, const int32_t *grid, int32_t *filtered, void *stream);
/* jfa.py:703-762 for every foam (rows are batch particles).  check: read
 * the row-table overflow flag on the host and grow the tables and emit again
 * (synchronises); 0: no host round trip, the caller reads the flag later
 * (fos_batch_flags) -- for frames captured as a graph. */
int fos_batch_face_accumulate(fos_batch *batch, fos_ctx *ctx, const int32_t *filtered,
                              int32_t *face_ids, int32_t *face_counts, int check, void *stream);
/* jfa.py:789-857 for every foam (each foam's voxel edge in the scores).
 * fixed_pairs > 0: that many fixed-point round pairs (2 rounds each) with no
 * host round trip (a device flag marks convergence, fos_batch_flags); 0:
 * until converged, the host checking after each burst of rounds. */
int fos_batch_finalize_faces(fos_batch *batch, fos_ctx *ctx, const float *rad,
                             const int32_t *face_ids, const int32_t *face_counts,
                             float *prev_scores, int32_t *neighbor_list, int32_t *neighbor_count,
                             int32_t *used_count, int32_t *fsc, int fixed_pairs, void *stream);
/* Host out[2] = {a row table overflowed in an emission, a fixed-round
 * finalize ended before its fixed point}; both stay set once set (the
 * context's flags), so one read covers every frame since the context was
 * made.  Synchronises the stream. */
int fos_batch_flags(fos_ctx *ctx, int32_t *out, void *stream);
/* grid.py:30-309 for every foam into one CSR over the batch's cells
 * (cell_count/cell_start [total cells], cell_indices [total particles],
 * batch-global).  With passes given (device int[B]), foams with
 * pass >= passes[b] are left as they are (their rebuild is not due).  With
 * cell_pr given (and rad), cell_pr[total particles][4] receives each CSR
 * slot's (wrapped position, radius), which the batched PBD and max-overlap
 * kernels read. */
int fos_batch_grid_build(fos_batch *batch, fos_ctx *ctx, float *pos, float L,
                         int32_t *cell_count, int32_t *cell_start, int32_t *cell_indices,
                         const int32_t *passes, int pass, const float *rad, float *cell_pr,
                         void *stream);
/* dynamics.py:304-372 for every foam; max_depth[B] (device) per foam.
 * cell_pr (may be NULL): the slot records of the last grid rebuild. */
int fos_batch_max_overlap(fos_batch *batch, const float *pos, const float *rad, float L,
                          const int32_t *cell_start, const int32_t *cell_count,
                          const int32_t *cell_indices, const float *cell_pr,
                          float *local_max_depth, float *max_depth, void *stream);
/* dynamics.py:77-171 (Jacobi) pass `pass` for foams with pass < passes[b]
 * (device int[B]); the others copy their positions.  cell_pr (may be NULL):
 * the slot records of the grid rebuild of pos_in. */
int fos_batch_pbd_pass(fos_batch *batch, const float *pos_in, float *pos_out, const float *rad,
                       float L, float gap_factor, float max_disp_frac, float disp_mult,
                       const int32_t *cell_start, const int32_t *cell_count,
                       const int32_t *cell_indices, const float *cell_pr,
                       const int32_t *passes, int pass, void *stream);
/* dynamics.py:1366-1427 for foams with active[b] (device int[B]); back[b]
 * (device float[B]) is the global backpressure factor of mode 1. */
int fos_batch_nudge_radii(fos_batch *batch, float *rad, const float *rad_target,
                          const float *local_max_depth, float alpha, float max_step_pct,
                          int mode, float contact_tol, const float *back,
                          const int32_t *active, void *stream);
/* dynamics.py:1281-1320 for foams with active[b], each with its FSC band and
 * radius bounds. */
int fos_batch_fsc_targets(fos_batch *batch, const float *rad, float *rad_target,
                          const int32_t *fsc, const float *fsc_ema, float growth_pct,
                          float deadband, const int32_t *active, void *stream);
/* dynamics.py:1497-1557 x iters for every foam (each foam's radius bounds;
 * the rotation hash takes the index within the foam).  frame_dev (may be
 * NULL): the frame number in device memory, read at run time. */
int fos_batch_pressure_run(fos_batch *batch, fos_ctx *ctx, int frame, const int32_t *frame_dev,
                           int iters, float *rad, float *delta_r, const int32_t *face_ids,
                           const int32_t *fsc, float k, float p_exp, float pair_cap, int max_nei,
                           void *stream);
/* dynamics.py:1596-1634 for every foam, Philox keyed (foam seed, frame,
 * index within the foam); frame_dev as for fos_batch_pressure_run. */
int fos_batch_brownian(fos_batch *batch, float *pos, float strength, float L, int frame,
                       const int32_t *frame_dev, int mode, void *stream);
/* out[b] (device int[B]) = particles of foam b with fsc == 0 (the
 * warm-start gate, run.py:1184-1195). */
int fos_batch_count_zero(fos_batch *batch, const int32_t *fsc, int32_t *out, void *stream);
/* The host driver's per-foam frame decisions on the device (all arrays
 * device, [B]): from the overlap maxima, each foam's PBD pass count
 * (run.py:709-714, r_max in double as the driver holds it), its global
 * backpressure factor and whether it settles this frame (settle counters
 * count down; run.py:733-751). */
int fos_batch_control_pre(fos_batch *batch, const float *max_depth, const double *r_max,
                          int pbd_base, int pbd_max, double pbd_scale, double contact_tol,
                          int32_t *settle, int32_t *passes, int32_t *active, float *back,
                          void *stream);
/* After the JFA (ran = 1) or a frame without it (ran = 0): the warm-start
 * gate per foam from its zero-FSC count and the device frame number
 * (run.py:1184-1195); growing foams restart their settle counters. */
int fos_batch_control_post(fos_batch *batch, const int32_t *zeros, const int32_t *frame_dev,
                           int warmstart_frames, double fsc_zero_rate, int adjustment_frames,
                           int ran, int32_t *settle, int32_t *growth, void *stream);
/* *frame_dev += 1 on the device. */
int fos_batch_frame_advance(int32_t *frame_dev, void *stream);

#ifdef __cplusplus
}
#endif
#endif
