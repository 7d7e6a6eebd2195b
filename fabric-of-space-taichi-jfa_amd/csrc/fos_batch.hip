// fos_batch.hip -- batches of independent foams (no reference counterpart).
//
// A batch concatenates the particles, JFA labels and hash cells of B foams of
// uneven size.  The descriptors (fos_inst) and the work tables built here let
// each phase of the frame run as one launch over all foams: workgroups find
// their foam from their tile or cell-range index, threads from their
// particle's foam id.
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "fos_common.h"

using namespace fos;

extern "C" {

int fos_batch_create(fos_batch **out, int B, const int *n, const int *res, const int *passes,
                     const int *G, const float *r_min, const float *r_max, const float *f_low,
                     const float *f_high, const unsigned *seed, float L) {
  FOS_REQUIRE(out && B >= 1 && n && res && passes && G && r_min && r_max && f_low && f_high &&
                  seed && L > 0.0f,
              "fos_batch_create: bad arguments (B=%d)", B);
  *out = nullptr;
  std::vector<fos_inst> h(B);
  std::vector<int> ts(B + 1), es(B + 1), gs(B + 1);
  long long v = 0, p = 0, c = 0, t = 0, e = 0, g = 0;
  int max_res = 0, max_passes = 0, max_G = 0;
  for (int b = 0; b < B; ++b) {
    FOS_REQUIRE(n[b] >= 1 && res[b] >= 2 && passes[b] >= 0 &&
                    passes[b] <= FOS_BATCH_PASSES - 1 && G[b] >= 3,
                "fos_batch_create: foam %d: n=%d res=%d passes=%d G=%d", b, n[b], res[b],
                passes[b], G[b]);
    fos_inst &I = h[b];
    I.n = n[b];
    I.p0 = (int)p;
    I.res = res[b];
    I.passes = passes[b];
    I.v0 = v;
    I.voxel = (float)((double)L / (double)res[b]);
    I.G = G[b];
    I.c0 = (int)c;
    I.r_min = r_min[b];
    I.r_max = r_max[b];
    I.f_low = f_low[b];
    I.f_high = f_high[b];
    I.seed = seed[b];
    I.tile = 0;
    const long long r = res[b], C = (long long)G[b] * G[b] * G[b];
    const long long nyb = (r + 3) / 4, nzb = (r + 63) / 64;
    ts[b] = (int)t;
    es[b] = (int)e;
    gs[b] = (int)g;
    t += r * nyb * nzb;
    e += ((r + EMIT_B_TX - 1) / EMIT_B_TX) * ((r + 256 / EMIT_B_ZW - 1) / (256 / EMIT_B_ZW)) *
         ((r + EMIT_B_ZW - 1) / EMIT_B_ZW);
    g += (C + fos_range_cells((int)C) - 1) / fos_range_cells((int)C);
    v += r * r * r;
    p += n[b];
    c += C;
    max_res = res[b] > max_res ? res[b] : max_res;
    max_passes = passes[b] > max_passes ? passes[b] : max_passes;
    max_G = G[b] > max_G ? G[b] : max_G;
    FOS_REQUIRE(p < (1ll << 31) && c < (1ll << 31) && t < (1ll << 31) && e < (1ll << 31),
                "fos_batch_create: the batch exceeds 32-bit particle, cell or tile indices");
  }
  ts[B] = (int)t;
  es[B] = (int)e;
  gs[B] = (int)g;
  std::vector<int> pinst(p);
  std::vector<float> pvox(p);
  for (int b = 0; b < B; ++b)
    for (int i = 0; i < n[b]; ++i) {
      pinst[h[b].p0 + i] = b;
      pvox[h[b].p0 + i] = h[b].voxel;
    }
  fos_batch *bt = (fos_batch *)calloc(1, sizeof(fos_batch));
  FOS_REQUIRE(bt, "fos_batch_create: out of host memory");
  bt->B = B;
  bt->n_total = (int)p;
  bt->c_total = (int)c;
  bt->v_total = v;
  bt->max_res = max_res;
  bt->max_passes = max_passes;
  bt->max_G = max_G;
  bt->inst_h = (fos_inst *)malloc(sizeof(fos_inst) * B);
  bt->tstart_h = (int *)malloc(sizeof(int) * (B + 1));
  bt->mstart_h = (int *)malloc(sizeof(int) * (B + 1));
  bt->estart_h = (int *)malloc(sizeof(int) * (B + 1));
  bt->gstart_h = (int *)malloc(sizeof(int) * (B + 1));
  if (!bt->inst_h || !bt->tstart_h || !bt->mstart_h || !bt->estart_h || !bt->gstart_h) {
    fos_batch_destroy(bt);
    FOS_REQUIRE(false, "fos_batch_create: out of host memory");
  }
  memcpy(bt->inst_h, h.data(), sizeof(fos_inst) * B);
  memcpy(bt->tstart_h, ts.data(), sizeof(int) * (B + 1));
  memcpy(bt->mstart_h, ts.data(), sizeof(int) * (B + 1));
  memcpy(bt->estart_h, es.data(), sizeof(int) * (B + 1));
  memcpy(bt->gstart_h, gs.data(), sizeof(int) * (B + 1));
  hipError_t err = hipSuccess;
#define BALLOC(ptr, bytes)                                    \
  if (err == hipSuccess) err = hipMalloc((void **)&(ptr), (bytes))
  BALLOC(bt->inst, sizeof(fos_inst) * B);
  BALLOC(bt->tstart, sizeof(int) * (B + 1));
  BALLOC(bt->mstart, sizeof(int) * (B + 1));
  BALLOC(bt->estart, sizeof(int) * (B + 1));
  BALLOC(bt->gstart, sizeof(int) * (B + 1));
  BALLOC(bt->pinst, sizeof(int) * p);
  BALLOC(bt->pvox, sizeof(float) * p);
  BALLOC(bt->stripes, sizeof(int) * (size_t)B * FOS_BATCH_PASSES * FOS_STRIPE_PASS_INTS);
  BALLOC(bt->fred, sizeof(float) * (size_t)B * 8);
  BALLOC(bt->ired, sizeof(int) * (size_t)B * 8);
#undef BALLOC
  if (err == hipSuccess) err = hipMemcpy(bt->inst, h.data(), sizeof(fos_inst) * B, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(bt->tstart, ts.data(), sizeof(int) * (B + 1), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(bt->mstart, ts.data(), sizeof(int) * (B + 1), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(bt->estart, es.data(), sizeof(int) * (B + 1), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(bt->gstart, gs.data(), sizeof(int) * (B + 1), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(bt->pinst, pinst.data(), sizeof(int) * p, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(bt->pvox, pvox.data(), sizeof(float) * p, hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    fos_batch_destroy(bt);
    return hip_fail(err, "fos_batch_create");
  }
  *out = bt;
  return 0;
}

int fos_batch_destroy(fos_batch *bt) {
  if (!bt) return 0;
  hipFree(bt->inst);
  hipFree(bt->tstart);
  hipFree(bt->mstart);
  hipFree(bt->estart);
  hipFree(bt->gstart);
  hipFree(bt->pinst);
  hipFree(bt->pvox);
  hipFree(bt->stripes);
  hipFree(bt->fred);
  hipFree(bt->ired);
  free(bt->inst_h);
  free(bt->tstart_h);
  free(bt->mstart_h);
  free(bt->estart_h);
  free(bt->gstart_h);
  free(bt);
  return 0;
}

int fos_batch_set_tiles(fos_batch *bt, const int *kinds) {
  FOS_REQUIRE(bt && kinds, "fos_batch_set_tiles: bad arguments");
  const int B = bt->B;
  std::vector<int> ts(B + 1);
  long long t = 0;
  for (int b = 0; b < B; ++b) {
    FOS_REQUIRE(kinds[b] == 0 || kinds[b] == 1, "fos_batch_set_tiles: foam %d: tile %d", b,
                kinds[b]);
    const long long r = bt->inst_h[b].res;
    ts[b] = (int)t;
    t += kinds[b] ? r * ((r * r + 255) / 256) : r * ((r + 3) / 4) * ((r + 63) / 64);
    FOS_REQUIRE(t < (1ll << 31), "fos_batch_set_tiles: tile count exceeds 32 bits");
    bt->inst_h[b].tile = kinds[b];
  }
  ts[B] = (int)t;
  memcpy(bt->tstart_h, ts.data(), sizeof(int) * (B + 1));
  FOS_HIP(hipMemcpy(bt->tstart, ts.data(), sizeof(int) * (B + 1), hipMemcpyHostToDevice));
  FOS_HIP(hipMemcpy(bt->inst, bt->inst_h, sizeof(fos_inst) * B, hipMemcpyHostToDevice));
  return 0;
}

int fos_batch_info(const fos_batch *bt, int64_t *out) {
  FOS_REQUIRE(bt && out, "fos_batch_info: bad arguments");
  out[0] = bt->B;
  out[1] = bt->n_total;
  out[2] = bt->v_total;
  out[3] = bt->c_total;
  out[4] = bt->max_res;
  out[5] = bt->max_passes;
  out[6] = bt->max_G;
  return 0;
}

}  // extern "C"
