// fos_frame.hip -- one whole foam frame from the host side in C++: the
// run.py:700-1261 order of foam.FoamSim.step() (plain JFA path) issued as
// one library call instead of ~70 calls from Python.  It calls the entry
// points FoamSim.step() calls, in the same order, or their fused forms
// (fos_jfa_faces_controller for faces + EMA + gate + targets), with the
// same two host reads (the overlap maximum before the PBD loop, the face-row
// flags inside the face step), so a frame gives the same bits either way.
#include "fos_common.h"

namespace {

// run.py:1071-1078 cadence: every frame before jfa_warm, then every
// jfa_every-th frame (foam_config.FoamConfig.jfa_runs).
bool jfa_runs(const fos_frame_desc *d, int frame) {
  if (frame < d->jfa_warm) return true;
  const int every = d->jfa_every > 1 ? d->jfa_every : 1;
  return (frame - d->jfa_warm) % every == 0;
}

// The face step's arguments for frame `frame` (fos_jfa_faces_controller).
fos::FaceStep face_step(const fos_frame_desc *d, const fos_frame_bufs *b, int frame) {
  fos::FaceStep a;
  a.res = d->res;
  a.n = d->n;
  a.filtered = b->filtered;
  a.rad_target = b->rad_target;
  a.rad = b->rad;
  a.voxel = d->voxel;
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
  if (ctx->spec_pending) FOS_CALL(fos::face_step_finish(ctx, face_step(d, b, f - 1), st));
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
  // 8 JFA on cadence frames (jfa.run_jfa without the host copies)
  const bool ran = jfa_runs(d, f);
  if (ran) {
    FOS_CALL(fos_jfa_init(ctx, pos, b->rad, n, d->res, L, d->voxel, d->beta, b->grid, nullptr,
                          stream));
    FOS_CALL(fos_jfa_passes_timed(ctx, d->res, L, d->voxel, d->jfa_passes, d->early_exit,
                                  d->th_pass_delta, d->min_passes, b->grid, b->temp, b->info,
                                  pass_events, stream));
    FOS_CALL(fos_majority(d->res, b->grid, b->filtered, stream));
    // face rows, finalize and the FSC EMA, then 9: the warm-start gate and
    // the gated targets on the device
    if (defer)
      FOS_CALL(fos::face_step_defer(ctx, face_step(d, b, f), st));
    else
      FOS_CALL(fos_jfa_faces_controller(
          ctx, d->res, n, b->filtered, b->rad_target, b->rad, d->voxel, b->face_ids,
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
