// fos_frame.hip -- one whole foam frame from the host side in C++: the
// run.py:700-1261 order of foam.FoamSim.step() (plain JFA path) issued as
// one library call instead of ~70 calls from Python.  It calls the entry
// points FoamSim.step() calls, in the same order, or their fused forms
// (fos_jfa_faces_controller for faces + EMA + gate + targets), with the
// same two host reads (the overlap maximum before the PBD loop, the face-row
// flags inside the face step), so a frame gives the same bits either way.
#include "fos_common.h"

namespace fos {
// fos_dynamics.hip: a PBD pass over the slot records without the next
// rebuild's cell step
int pbd_pass_records(const float *pos_in, float *pos_out, const float *rad, int n, int G,
                     float L, float gap_factor, float max_disp_frac, float disp_mult,
                     const int32_t *cell_start, const int32_t *cell_count,
                     const int32_t *cell_indices, const float *cell_pr, hipStream_t st);
}  // namespace fos

namespace {

// run.py:1071-1078 cadence: every frame before jfa_warm, then every
// jfa_every-th frame (foam_config.FoamConfig.jfa_runs).
bool jfa_runs(const fos_frame_desc *d, int frame) {
  if (frame < d->jfa_warm) return true;
  const int every = d->jfa_every > 1 ? d->jfa_every : 1;
  return (frame - d->jfa_warm) % every == 0;
}

// jfa.py:224 pass count, ceil(log2 res) + 1
int jfa_passes_for(int res) { return (int)ceil(log2((double)res)) + 1; }

// The face step's arguments for frame `frame` (fos_jfa_faces_controller), at
// the resolution that frame's JFA ran at (s->jfa_res; 0: the configured one).
fos::FaceStep face_step(const fos_frame_desc *d, const fos_frame_bufs *b, int frame,
                        const fos_frame_state *s) {
  fos::FaceStep a;
  const int res = s->jfa_res > 0 ? s->jfa_res : d->res;
  a.res = res;
  a.n = d->n;
  a.filtered = b->filtered;
  a.rad_target = b->rad_target;
  a.rad = b->rad;
  a.voxel = res == d->res ? d->voxel : (float)(d->L_d / (double)res);
  a.face_ids = b->face_ids;
  a.face_counts = b->face_counts;
  a.prev_scores = b->prev_scores;
  a.neighbor_list = b->neighbor_list;
  a.neighbor_count = b->neighbor_count;
  a.used_count = b->used_count;
  a.fsc = b->fsc;
  a.fsc_ema = b->fsc_ema;
  a.alpha = d->ema_alpha;
  a.frame = frame;
  a.warmstart_frames = d->warmstart_frames;
  a.fsc_zero_rate = d->fsc_zero_rate;
  a.f_low = d->f_low;
  a.f_high = d->f_high;
  a.growth_pct = d->growth_pct;
  a.r_min = d->r_min;
  a.r_max = d->r_max;
  a.deadband = d->deadband;
  a.gate = b->maxd + 1;
  return a;
}

}  // namespace

// ---------------------------------------------------------------------------
// Temporal-skip scheduler statistics (run.py:902-1018), in a fixed order:
// stage A, per block, sums |dr|, r, fsc, the out-of-band count and the
// maximum squared displacement; stage B (one block) totals them; stage C,
// per block, sums (|dr| - mean)^2 and moves the current values into the
// previous-value buffers; stage D (one block) writes the result.
// ---------------------------------------------------------------------------
#define SCHED_T 256

// Block sum of 8 doubles in a fixed tree order; thread 0 gets the result.
__device__ __forceinline__ void sched_block_sum(double v[8], bool max5, double *dst) {
  __shared__ double s_v[8][SCHED_T];
  const int t = (int)threadIdx.x;
#pragma unroll
  for (int k = 0; k < 8; ++k) s_v[k][t] = v[k];
  __syncthreads();
  for (int o = SCHED_T / 2; o > 0; o >>= 1) {
    if (t < o) {
#pragma unroll
      for (int k = 0; k < 8; ++k)
        s_v[k][t] = (max5 && k == 5) ? fmax(s_v[k][t], s_v[k][t + o]) : s_v[k][t] + s_v[k][t + o];
    }
    __syncthreads();
  }
  if (t == 0) {
#pragma unroll
    for (int k = 0; k < 8; ++k) dst[k] = s_v[k][0];
  }
}

// |pos - pos_prev| of particle i at the minimum image, as numpy computes it
// in fp32: d = p - pp; |d| > L/2 -> d - sign(d) L; sqrt((dx^2 + dy^2) + dz^2).
__device__ __forceinline__ float sched_disp(const float *__restrict__ pos,
                                            const float *__restrict__ pp, int i, float L) {
  float s = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float d = pos[3 * i + a] - pp[3 * i + a];
    if (fabsf(d) > L * 0.5f) d = d - copysignf(L, d);
    s = s + d * d;
  }
  return sqrtf(s);
}

__global__ void __launch_bounds__(SCHED_T)
    k_sched_a(const float *__restrict__ pos, const float *__restrict__ rad,
              const int32_t *__restrict__ fsc, int n, float L, int f_low, int f_high,
              const float *__restrict__ pp, const float *__restrict__ rp, int have_prev,
              double *__restrict__ part) {
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = blockIdx.x * SCHED_T + threadIdx.x; i < n; i += gridDim.x * SCHED_T) {
    const float r = rad[i];
    const int f = fsc[i];
    if (have_prev) {
      v[0] += (double)fabsf(r - rp[i]);
      v[5] = fmax(v[5], (double)sched_disp(pos, pp, i, L));
    }
    v[1] += (double)r;
    v[2] += (double)f;
    v[3] += (f < f_low || f > f_high) ? 1.0 : 0.0;
  }
  sched_block_sum(v, true, part + 8 * blockIdx.x);
}

// Stage B and D: one block totals the partials of `blocks` blocks (slot 5
// a maximum in stage B).
__global__ void __launch_bounds__(SCHED_T)
    k_sched_total(const double *__restrict__ part, int blocks, bool max5,
                  double *__restrict__ dst) {
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int t = (int)threadIdx.x;
  if (t < blocks) {
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = part[8 * t + k];
  }
  sched_block_sum(v, max5, dst);
}

__global__ void __launch_bounds__(SCHED_T)
    k_sched_c(const float *__restrict__ pos, const float *__restrict__ rad, int n,
              float *__restrict__ pp, float *__restrict__ rp, int have_prev,
              const double *__restrict__ tot, double *__restrict__ part) {
  const double mean = have_prev ? tot[0] / (double)n : 0.0;
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = blockIdx.x * SCHED_T + threadIdx.x; i < n; i += gridDim.x * SCHED_T) {
    const float r = rad[i];
    if (have_prev) {
      const double e = (double)fabsf(r - rp[i]) - mean;
      v[0] += e * e;
    }
    rp[i] = r;
    pp[3 * i] = pos[3 * i];
    pp[3 * i + 1] = pos[3 * i + 1];
    pp[3 * i + 2] = pos[3 * i + 2];
  }
  sched_block_sum(v, false, part + 8 * blockIdx.x);
}

__global__ void k_sched_out(const double *__restrict__ tot, const double *__restrict__ tot2,
                            int n, double *__restrict__ out) {
  const double dn = (double)n;
  out[0] = tot[0] / dn;
  out[1] = sqrt(tot2[0] / dn);
  out[2] = tot[5];
  out[3] = tot[1] / dn;
  out[4] = tot[2] / dn;
  out[5] = tot[3];
  out[6] = 0.0;
  out[7] = 0.0;
}

#define FOS_CALL(x)             \
  do {                          \
    const int _rc = (x);        \
    if (_rc) return _rc;        \
  } while (0)

// One frame.  defer: the face step goes without its host read, which the
// next frame makes once it has queued the work that does not read the faces
// (the rebuild and the overlap maximum).
static int frame_step(fos_ctx *ctx, const fos_frame_desc *d, const fos_frame_bufs *b,
                      fos_frame_state *s, void *const *pass_events, void *stream, bool defer) {
  FOS_REQUIRE(ctx && d && b && s && d->n >= 1 && d->n <= ctx->max_n && d->res >= 1 &&
                  d->res <= ctx->max_res && d->G >= 3 &&
                  (int64_t)d->G * d->G * d->G <= ctx->max_cells && b->maxd_host &&
                  (s->pos_cur == 0 || s->pos_cur == 1) && (s->sidx == 0 || s->sidx == 1),
              "fos_frame_step: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const int n = d->n, G = d->G, f = s->frame;
  const float L = d->L;
  float *pos = b->pos[s->pos_cur], *pos_tmp = b->pos[1 - s->pos_cur];
  if (!s->event) {
    hipEvent_t ev;
    FOS_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    s->event = (void *)ev;
  }
  // 0-1: the rebuild at the end of the last frame built this CSR from these
  // positions unless the caller wrote them since (need_rebuild); with L = 1
  // the wrap is idempotent, so rebuilding again reproduces it bit for bit
  if (L != 1.0f || s->need_rebuild)
    FOS_CALL(fos_grid_build(ctx, pos, n, G, L, b->cell_count, b->cell_start, b->cell_indices,
                            stream));
  FOS_CALL(fos_max_overlap(pos, b->rad, n, G, L, b->cell_start, b->cell_count, b->cell_indices,
                           b->lmd, b->maxd, stream));
  // the last frame's deferred face step is finished before anything reads
  // its faces, FSC, targets or gate
  if (ctx->spec_pending) FOS_CALL(fos::face_step_finish(ctx, face_step(d, b, f - 1, s), st));
  // the overlap maximum (and the last gated frame's gate) go to the host
  // behind an event; settle, pressure and noise need neither and are queued
  // before the host waits
  FOS_HIP(hipMemcpyAsync(b->maxd_host, b->maxd, 2 * sizeof(float), hipMemcpyDeviceToHost, st));
  FOS_HIP(hipEventRecord((hipEvent_t)s->event, st));
  // 2 settle (counter and gate on the device, two slots used in turn)
  FOS_CALL(fos_settle_nudge(b->rad, b->rad_target, b->lmd, n, d->alpha, d->max_step_pct,
                            d->r_min, d->r_max, d->backpressure_mode, d->contact_tol,
                            d->contact_tol_d, b->maxd, b->settle_dev + s->sidx,
                            b->settle_dev + (1 - s->sidx), d->adjustment_frames,
                            s->settle_force, s->settle_gate, stream));
  s->sidx = 1 - s->sidx;
  s->settle_force = s->settle_gate = 0;
  // 3 pressure
  if (d->pressure)
    FOS_CALL(fos_pressure_run(ctx, n, f, d->pressure_iters, b->rad, b->delta_r, b->face_ids,
                              b->fsc, d->pressure_k, d->pressure_exp, d->pressure_cap,
                              d->pressure_max_nei, d->r_min, d->r_max, nullptr, stream));
  // 4 Brownian; with PBD passes to follow it also runs the wrap-and-hash step
  // of the first pass's rebuild
  bool ready = false;
  if (d->brownian) {
    if (d->pbd_base > 0) {
      FOS_CALL(fos_brownian_cells(ctx, pos, n, d->brownian_strength, L, d->seed, f,
                                  d->brownian_mode, G, stream));
      ready = true;
    } else {
      FOS_CALL(fos_brownian(pos, n, d->brownian_strength, L, d->seed, f, d->brownian_mode,
                            stream));
    }
  }
  FOS_HIP(hipEventSynchronize((hipEvent_t)s->event));
  const float max_depth = b->maxd_host[0];
  s->gate_read = s->pending;
  if (s->pending) s->gate = b->maxd_host[1];
  s->pending = 0;
  // run.py:712-714 adaptive pass count (Python float arithmetic: double)
  const int raw = (int)((double)d->pbd_base + d->pbd_scale * ((double)max_depth /
                                                              (0.2 * d->r_max_d)));
  int passes = raw < d->pbd_max ? raw : d->pbd_max;
  if (passes < d->pbd_base) passes = d->pbd_base;
  float *pr = d->slot_records ? b->cell_pr : nullptr;
  // 5 PBD; a pass followed directly by a rebuild runs that rebuild's
  // wrap-and-hash step itself
  for (int p = 0; p < passes; ++p) {
    if (ready)
      FOS_CALL(fos_grid_build_from_cells(ctx, pos, b->rad, n, G, b->cell_count, b->cell_start,
                                         b->cell_indices, pr, stream));
    else
      FOS_CALL(fos_grid_build_pr(ctx, pos, b->rad, n, G, L, b->cell_count, b->cell_start,
                                 b->cell_indices, pr, stream));
    ready = p < passes - 1 || !d->ou_jitter;
    if (ready)
      FOS_CALL(fos_pbd_pass_cells(ctx, pos, pos_tmp, b->rad, n, G, L, d->gap_factor,
                                  d->max_disp_frac, d->disp_mult, b->cell_start, b->cell_count,
                                  b->cell_indices, pr, stream));
    else if (pr)
      FOS_CALL(fos::pbd_pass_records(pos, pos_tmp, b->rad, n, G, L, d->gap_factor,
                                     d->max_disp_frac, d->disp_mult, b->cell_start,
                                     b->cell_count, b->cell_indices, pr, st));
    else
      FOS_CALL(fos_pbd_pass(pos, pos_tmp, b->rad, n, G, L, d->gap_factor, d->max_disp_frac,
                            d->disp_mult, b->cell_start, b->cell_count, b->cell_indices,
                            stream));
    float *t = pos;
    pos = pos_tmp;
    pos_tmp = t;
    s->pos_cur = 1 - s->pos_cur;
  }
  // 6 OU jitter (mean radius on the device)
  if (d->ou_jitter)
    FOS_CALL(fos_ou_jitter_mean(ctx, pos, b->v_jit, b->rad, n, d->dt, d->jitter_rms,
                                d->jitter_tau, d->drift_cap_frac, L, d->seed, f, nullptr,
                                stream));
  // 7
  if (ready)
    FOS_CALL(fos_grid_build_from_cells(ctx, pos, b->rad, n, G, b->cell_count, b->cell_start,
                                       b->cell_indices, nullptr, stream));
  else
    FOS_CALL(fos_grid_build(ctx, pos, n, G, L, b->cell_count, b->cell_start, b->cell_indices,
                            stream));
  s->need_rebuild = 0;
  // 8 JFA on the frames the schedule picks (jfa.run_jfa without the host
  // copies), at the resolution it picks
  int run = 0, res = d->res;
  FOS_CALL(fos_jfa_schedule(ctx, d, b, s, pos, &run, &res, stream));
  const bool ran = run != 0;
  if (ran) {
    const float voxel = res == d->res ? d->voxel : (float)(d->L_d / (double)res);
    const int passes = res == d->res ? d->jfa_passes : jfa_passes_for(res);
    s->jfa_res = res;
    FOS_CALL(fos_jfa_init(ctx, pos, b->rad, n, res, L, voxel, d->beta, b->grid, nullptr,
                          stream));
    FOS_CALL(fos_jfa_passes_timed(ctx, res, L, voxel, passes, d->early_exit,
                                  d->th_pass_delta, d->min_passes, b->grid, b->temp, b->info,
                                  pass_events, stream));
    FOS_CALL(fos_majority(res, b->grid, b->filtered, stream));
    // face rows, finalize and the FSC EMA, then 9: the warm-start gate and
    // the gated targets on the device
    if (defer)
      FOS_CALL(fos::face_step_defer(ctx, face_step(d, b, f, s), st));
    else
      FOS_CALL(fos_jfa_faces_controller(
          ctx, res, n, b->filtered, b->rad_target, b->rad, voxel, b->face_ids,
          b->face_counts, b->prev_scores, b->neighbor_list, b->neighbor_count, b->used_count,
          b->fsc, b->fsc_ema, d->ema_alpha, f, d->warmstart_frames, d->fsc_zero_rate, d->f_low,
          d->f_high, d->growth_pct, d->r_min, d->r_max, d->deadband, b->maxd + 1, stream));
    s->pending = 1;
    s->settle_gate = 1;
  } else {
    FOS_CALL(fos_fsc_targets(b->rad, b->rad_target, b->fsc, b->fsc_ema, n, d->f_low, d->f_high,
                             d->growth_pct, d->r_min, d->r_max, d->deadband, stream));
    s->settle_force = 1;
  }
  s->pbd_passes = passes;
  s->max_depth = max_depth;
  s->jfa_ran = ran ? 1 : 0;
  s->frame = f + 1;
  return 0;
}

extern "C" {

int fos_sched_stats(fos_ctx *ctx, const float *pos, const float *rad, const int32_t *fsc, int n,
                    float L, int f_low, int f_high, float *pos_prev, float *rad_prev,
                    int have_prev, double *out, void *stream) {
  FOS_REQUIRE(ctx && pos && rad && fsc && pos_prev && rad_prev && out && n >= 1 &&
                  n <= ctx->max_n,
              "fos_sched_stats: bad arguments (n=%d)", n);
  hipStream_t st = (hipStream_t)stream;
  int blocks = (n + SCHED_T - 1) / SCHED_T;
  if (blocks > FOS_SCHED_BLOCKS) blocks = FOS_SCHED_BLOCKS;
  double *part = ctx->sched, *tot = ctx->sched + 8 * FOS_SCHED_BLOCKS, *tot2 = tot + 8;
  hipLaunchKernelGGL(k_sched_a, dim3(blocks), dim3(SCHED_T), 0, st, pos, rad, fsc, n, L, f_low,
                     f_high, pos_prev, rad_prev, have_prev, part);
  hipLaunchKernelGGL(k_sched_total, dim3(1), dim3(SCHED_T), 0, st, part, blocks, true, tot);
  hipLaunchKernelGGL(k_sched_c, dim3(blocks), dim3(SCHED_T), 0, st, pos, rad, n, pos_prev,
                     rad_prev, have_prev, tot, part);
  hipLaunchKernelGGL(k_sched_total, dim3(1), dim3(SCHED_T), 0, st, part, blocks, false, tot2);
  hipLaunchKernelGGL(k_sched_out, dim3(1), dim3(1), 0, st, tot, tot2, n, out);
  FOS_CHECK_LAUNCH("k_sched");
  return 0;
}

int fos_jfa_schedule(fos_ctx *ctx, const fos_frame_desc *d, const fos_frame_bufs *b,
                     fos_frame_state *s, const float *pos, int *run, int *res, void *stream) {
  FOS_REQUIRE(ctx && d && b && s && pos && run && res, "fos_jfa_schedule: bad arguments");
  *res = d->res;
  const bool temporal = d->schedule == 1;
  if (!temporal) {
    *run = jfa_runs(d, s->frame) ? 1 : 0;
    s->reason = *run ? FOS_JFA_CADENCE : FOS_JFA_SKIP;
  }
  if (!temporal && !(d->adaptive_res && *run)) return 0;
  FOS_REQUIRE(b->pos_prev && b->rad_prev && b->sched_host,
              "fos_jfa_schedule: the scheduler needs pos_prev, rad_prev and sched_host");
  hipStream_t st = (hipStream_t)stream;
  double *out = ctx->sched + 8 * (FOS_SCHED_BLOCKS + 2);
  FOS_CALL(fos_sched_stats(ctx, pos, b->rad, b->fsc, d->n, d->L, d->f_low, d->f_high,
                           b->pos_prev, b->rad_prev, temporal ? s->have_prev : 0, out, stream));
  FOS_HIP(hipMemcpyAsync(b->sched_host, out, 8 * sizeof(double), hipMemcpyDeviceToHost, st));
  FOS_HIP(hipStreamSynchronize(st));
  const double *o = b->sched_host;
  const double r_mean = o[3];
  if (temporal) {
    // run.py:906-1018
    s->sched_frames += 1;
    s->since += 1;
    double sig = 999.0, vmax = 999.0;
    if (s->have_prev) {
      sig = o[1] / (r_mean > 1e-9 ? r_mean : 1e-9);
      vmax = o[2];
    }
    s->have_prev = 1;
    // the FSC mean after the last JFA: the FSC has not changed since (only a
    // JFA writes it), so it is the mean measured now
    if (s->fsc_mu_pending) {
      s->last_fsc_mu = o[4];
      s->fsc_mu_set = 1;
      s->fsc_mu_pending = 0;
    }
    const double dmu = s->fsc_mu_set ? fabs(o[4] - s->last_fsc_mu) : 999.0;
    const double th_margin = d->th_margin * r_mean;
    if (s->lockout > 0) s->lockout -= 1;
    const bool warm = s->sched_frames <= d->skip_warmstart;
    const bool watchdog = s->since >= d->max_skip;
    const bool band = s->lockout > 0;
    const bool activity = sig > d->th_rsigma || vmax > th_margin || dmu > d->th_fsc_drift;
    const double oob = o[5] / (double)(d->n > 1 ? d->n : 1);
    const bool out_of_band = s->since > 1 && oob > d->oob_frac;
    s->signals[0] = sig;
    s->signals[1] = vmax;
    s->signals[2] = dmu;
    s->signals[3] = oob;
    s->signals[4] = r_mean;
    *run = (warm || watchdog || band || activity || out_of_band) ? 1 : 0;
    s->reason = warm ? FOS_JFA_WARMSTART
                : band ? FOS_JFA_BAND_CHANGE
                : watchdog ? FOS_JFA_WATCHDOG
                : out_of_band ? FOS_JFA_OUT_OF_BAND
                : activity ? FOS_JFA_ACTIVITY
                           : FOS_JFA_SKIP;
    if (*run) {
      s->since = 0;
      s->fsc_mu_pending = 1;
    }
  }
  if (*run && d->adaptive_res) {
    // run.py:1054-1061: round() of a Python float (ties to even)
    int r = (int)nearbyint(d->L_d * d->voxels_per_diameter / (2.0 * r_mean));
    r = r < d->res_min ? d->res_min : r;
    r = r > d->res_max ? d->res_max : r;
    FOS_REQUIRE(r >= 1 && r <= ctx->max_res,
                "fos_jfa_schedule: adaptive resolution %d outside the context's 1..%d", r,
                ctx->max_res);
    *res = r;
  }
  return 0;
}

int fos_frame_step(fos_ctx *ctx, const fos_frame_desc *d, const fos_frame_bufs *b,
                   fos_frame_state *s, void *const *pass_events, void *stream) {
  return frame_step(ctx, d, b, s, pass_events, stream, false);
}

int fos_frame_run(fos_ctx *ctx, const fos_frame_desc *d, const fos_frame_bufs *b,
                  fos_frame_state *s, int frames, void *const *const *pass_events,
                  void *stream) {
  FOS_REQUIRE(ctx && frames >= 0 && s, "fos_frame_run: bad arguments (frames=%d)", frames);
  // every frame but the last leaves its face step's host read to the next
  for (int k = 0; k < frames; ++k)
    FOS_CALL(frame_step(ctx, d, b, s, pass_events ? pass_events[k] : nullptr, stream,
                        k + 1 < frames));
  return 0;
}

int fos_frame_release(fos_frame_state *s) {
  if (s && s->event) {
    (void)hipEventDestroy((hipEvent_t)s->event);
    s->event = nullptr;
  }
  return 0;
}

}  // extern "C"
