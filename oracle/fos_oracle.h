/*
 * fos_oracle.h -- CPU restatement of the reference per-frame foam step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path and the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never links it.
 *
 * Every function restates the *serial* semantics of one reference Taichi
 * kernel (outermost loops executed by one thread in program order).  The
 * per-element loops whose iterations are independent may run under OpenMP;
 * the order-defined loops (face accumulation, finalize, grid scatter) run
 * serially.  Built with -ffp-contract=off and no fast-math, so each fp32
 * expression rounds exactly as written.
 *
 * Arrays: pos float[n*3] (AoS xyz), rad float[n], labels int32[res^3] with
 * index x*res*res + y*res + z, per-particle rows int32/float[n*32].
 */
#ifndef FOS_ORACLE_H
#define FOS_ORACLE_H
#include <stdint.h>

#define ORC_MAX_NEIGHBORS 32

#ifdef __cplusplus
extern "C" {
#endif

void orc_set_threads(int nthreads);
int orc_get_threads(void);

/* jfa.py:247-308 helpers */
float orc_power_distance(float cx, float cy, float cz, const float *p, float r,
                         float beta, float L);
int orc_seed_voxel(const float *p, int res, float L);

/* jfa.py:485-520 (jfa_init): clear grid+temp over res^3, seed particles. */
void orc_jfa_init(const float *pos, const float *rad, int n, int res, float L,
                  float voxel, float beta, int32_t *grid, int32_t *temp);
/* jfa.py:524-623 (jfa_step_ping0/1): one pass in -> out, returns changes. */
int64_t orc_jfa_step(const float *pos, const float *rad, int res, float L,
                     float voxel, float beta, int step, const int32_t *in,
                     int32_t *out);
/* jfa.py:1078-1116 (run_jfa pass loop).  Result ends in grid.
 * changes_hist[num_passes] receives per-pass changed-voxel counts.
 * Returns passes executed; *early_exit set to 1 when the exit fired. */
int orc_jfa_run(const float *pos, const float *rad, int n, int res, float L,
                float voxel, float beta, int num_passes, int early_exit_enabled,
                double th_pass_delta, int min_passes, int32_t *grid,
                int32_t *temp, int64_t *changes_hist, int *early_exit);
/* jfa.py:640-699 */
void orc_majority(const int32_t *grid, int32_t *filtered, int res);
/* jfa.py:703-762, serial slot scan */
void orc_accumulate_faces(const int32_t *filtered, int res, int n,
                          int32_t *face_ids, int32_t *face_counts);
/* jfa.py:379-443,789-857 */
void orc_finalize_faces(int n, const float *rad, float voxel,
                        const int32_t *face_ids, const int32_t *face_counts,
                        float *prev_scores, int32_t *neighbor_list,
                        int32_t *neighbor_count, int32_t *used_count,
                        int32_t *fsc);
/* jfa.py:766-785 */
void orc_fsc_ema(int n, float alpha, const int32_t *fsc, float *fsc_ema);
/* jfa.py:1204-1330: out_i[0]=p_dir, [1]=overflow, [2]=min_fsc, [3]=max_fsc,
 * [4]=sum_fsc; out_f[0]=mean_fsc, [1]=mean_ema, [2]=mean_score,
 * [3]=std_score. */
void orc_jfa_stats(int n, const float *rad, float voxel, const int32_t *fsc,
                   const float *fsc_ema, const int32_t *used_count,
                   const int32_t *face_ids, const int32_t *face_counts,
                   int64_t *out_i, double *out_f);
/* jfa.py:1337-1397: out[0]=asym, out[1]=overflow, out[2]=self_loops */
void orc_validate(int n, const int32_t *neighbor_list,
                  const int32_t *neighbor_count, int64_t *out);
/* jfa.py:967-1049 dirty tiles */
void orc_mark_dirty_tiles(const float *pos, const float *rad,
                          const float *pos_prev, const float *rad_prev, int n,
                          int res, float L, float voxel, int tile, int halo,
                          float pos_thr, float rad_thr, uint8_t *tile_dirty);

/* grid.py:145-309 rebuild: count, exclusive scan, stable scatter (stores
 * wrapped positions back into pos). */
void orc_grid_build(float *pos, int n, int G, float L, int32_t *cell_count,
                    int32_t *cell_start, int32_t *cell_indices);
/* dynamics.py:77-171, Jacobi form: reads pos_in, writes pos_out. */
void orc_pbd_jacobi(const float *pos_in, float *pos_out, const float *rad,
                    int n, int G, float L, float gap_factor, float max_disp_frac,
                    float disp_mult, const int32_t *cell_start,
                    const int32_t *cell_count, const int32_t *cell_indices);
/* dynamics.py:77-171, literal index-order Gauss-Seidel (serial). */
void orc_pbd_gauss_seidel(float *pos, const float *rad, int n, int G, float L,
                          float gap_factor, float max_disp_frac, float disp_mult,
                          const int32_t *cell_start, const int32_t *cell_count,
                          const int32_t *cell_indices);
/* dynamics.py:304-372 */
float orc_max_overlap(const float *pos, const float *rad, int n, int G, float L,
                      const int32_t *cell_start, const int32_t *cell_count,
                      const int32_t *cell_indices, float *local_max_depth);
/* dynamics.py:1242-1320 */
void orc_fsc_targets(const float *rad, float *rad_target, const int32_t *fsc,
                     const float *fsc_ema, int n, int f_low, int f_high,
                     float growth_pct, float r_min, float r_max,
                     float deadband);
/* dynamics.py:1322-1363: out_i[0..2]=grow,shrink,idle; out_f[0..1]=mean,max */
void orc_controller_stats(const float *rad, const float *rad_target,
                          const float *fsc_ema, int n, int f_low, int f_high,
                          int64_t *out_i, double *out_f);
/* dynamics.py:1366-1427 */
void orc_nudge_radii(float *rad, const float *rad_target,
                     const float *local_max_depth, int n, float alpha,
                     float max_step_pct, float r_min, float r_max,
                     float back_global, int mode, float contact_tol);
/* dynamics.py:1436-1451 */
int32_t orc_hash(int32_t i, int32_t frame);
/* dynamics.py:1454-1557; returns max|dr|, *changed */
float orc_pressure(int n, int frame, float *rad, float *delta_r,
                   const int32_t *face_ids, const int32_t *fsc, float k,
                   float p_exp, float pair_cap, int max_nei, float r_min,
                   float r_max, int32_t *changed);
/* Counter-based noise shared with the product: Philox4x32-10. */
void orc_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                uint32_t k0, uint32_t k1, uint32_t out[4]);
/* dynamics.py:1596-1634 with Philox noise.  mode 0: uniform
 * (u-0.5)*2*strength; mode 1: Gaussian sigma=strength (Box-Muller). */
void orc_brownian(float *pos, int n, float strength, float L, uint64_t seed,
                  int frame, int mode);
/* dynamics.py:607-660 OU jitter + integrate, Philox uniforms */
void orc_ou_jitter(float *pos, float *v_jit, const float *rad, int n,
                   float mean_r, float dt, float jitter_rms, float jitter_tau,
                   float drift_cap_frac, float L,
                   uint64_t seed, int frame);
/* dynamics.py:532-593 */
void orc_xsph(const float *pos, float *vel, float *vel_tmp, int n, int G,
              float L, float lim_sq, float eps, const int32_t *cell_start,
              const int32_t *cell_count, const int32_t *cell_indices);
/* topology.py:68-133: pairs[n*cap], count[n], truncated[n] */
void orc_topo_candidates(const float *pos, const float *rad, int n, int G,
                         float L, float mult, float r_max, int cap,
                         const int32_t *cell_start, const int32_t *cell_count,
                         const int32_t *cell_indices, int32_t *pairs,
                         int32_t *count, int32_t *truncated);

/* topology.py:140-504 */
void orc_topo_gabriel(const float *pos, int n, int G, float L, float cell_size,
                      const int32_t *cs, const int32_t *cc, const int32_t *ci,
                      const int32_t *pairs, const int32_t *pair_count, int cap, int frame,
                      int stride, int early_exit, float deg_high, const float *deg_ema,
                      int32_t *deg);
void orc_topo_gabriel_batched(const float *pos, const float *rad, int n, int G, float L,
                              float cell_size, float mult, int cap, const int32_t *cs,
                              const int32_t *cc, const int32_t *ci, int batch_idx,
                              int num_batches, int q, int frame, int32_t *pairs,
                              int32_t *pair_count, int32_t *truncated, int32_t *deg);
void orc_topo_knn(const float *pos, const float *rad, int n, int G, float L, float mult, int k,
                  const int32_t *cs, const int32_t *cc, const int32_t *ci, int32_t *deg);
void orc_topo_ema(const int32_t *deg, float *ema, int n, float alpha);
void orc_topo_radii(float *rad, const float *deg_eff, int n, float deg_low, float deg_high,
                    float gain_grow, float gain_shrink, float r_min, float r_max);

#ifdef __cplusplus
}
#endif
#endif
