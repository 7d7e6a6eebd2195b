// fos_topology.hip -- topological degree kernels of topology.py (Gabriel
// witness test, batched Gabriel, k-NN proxy, degree EMA, radius rule).
// They are off by default in the reference (TOPO_ENABLED = False,
// config.py:453) and sit beside the foam step, not on its per-frame path.
#include "fos_common.h"

using namespace fos;

__device__ __forceinline__ float d3(float x, float y, float z) { return (x * x + y * y) + z * z; }

// topology.py:188 and :326-333 pair hash (int32 arithmetic wraps).
__device__ __forceinline__ int topo_hash(int i, int j, int frame) {
  return (int)(((unsigned)i * 73856093u) ^ ((unsigned)j * 19349663u) ^
               ((unsigned)frame * 83492791u));
}

// topology.py:198-252: 1 when no particle other than i and j lies strictly
// inside the diameter sphere of (i, j) (scaled epsilon), 0 when one does,
// -1 for a degenerate pair.  The sphere is searched over an adaptive stencil
// of 1..3 cells (at most G/2) around the midpoint's cell.
__device__ int gabriel_pair(const float *__restrict__ pos, int i, int j, int G, float L,
                            float half_l, float inv_l, float cell_size,
                            const int32_t *__restrict__ cs, const int32_t *__restrict__ cc,
                            const int32_t *__restrict__ ci) {
  const float pix = pos[3 * i], piy = pos[3 * i + 1], piz = pos[3 * i + 2];
  const float dx = gwrap(pos[3 * j] - pix, L, inv_l), dy = gwrap(pos[3 * j + 1] - piy, L, inv_l),
              dz = gwrap(pos[3 * j + 2] - piz, L, inv_l);
  const float d2 = d3(dx, dy, dz);
  if (d2 < 1e-16f) return -1;
  const float mx = gwrap(pix + 0.5f * dx, L, inv_l), my = gwrap(piy + 0.5f * dy, L, inv_l),
              mz = gwrap(piz + 0.5f * dz, L, inv_l);
  const float r_sq = 0.25f * d2;
  const int cx = wrap_fast((int)(((mx + half_l) * inv_l) * (float)G), G);
  const int cy = wrap_fast((int)(((my + half_l) * inv_l) * (float)G), G);
  const int cz = wrap_fast((int)(((mz + half_l) * inv_l) * (float)G), G);
  int sc = (int)ceilf((sqrtf(d2) * 0.5f) / cell_size);
  sc = sc < 1 ? 1 : (sc > 3 ? 3 : sc);
  if (sc > G / 2) sc = G / 2;
  const float eps = fmaxf(1e-12f, 1e-6f * r_sq);
  for (int ax = -sc; ax <= sc; ++ax)
    for (int ay = -sc; ay <= sc; ++ay)
      for (int az = -sc; az <= sc; ++az) {
        const int id = (wrap_fast(cx + ax, G) * G + wrap_fast(cy + ay, G)) * G + wrap_fast(cz + az, G);
        const int b = cs[id], e = b + cc[id];
        for (int l = b; l < e; ++l) {
          const int k = ci[l];
          if (k == i || k == j) continue;
          const float ex = gwrap(mx - pos[3 * k], L, inv_l), ey = gwrap(my - pos[3 * k + 1], L, inv_l),
                      ez = gwrap(mz - pos[3 * k + 2], L, inv_l);
          if (d3(ex, ey, ez) < r_sq - eps) return 0;
        }
      }
  return 1;
}

// topology.py:140-262 gabriel_test_topological_degree: one thread per
// candidate slot; degrees are integer atomics (order-free).
__global__ void k_topo_gabriel(const float *__restrict__ pos, int n, int G, float L, float half_l,
                               float inv_l, float cell_size, const int32_t *__restrict__ cs,
                               const int32_t *__restrict__ cc, const int32_t *__restrict__ ci,
                               const int32_t *__restrict__ pairs,
                               const int32_t *__restrict__ pair_count, int cap, int frame,
                               int stride, int early_exit, float deg_high,
                               const float *__restrict__ deg_ema, int32_t *__restrict__ deg) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * cap) return;
  const int i = (int)(t / cap), p = (int)(t - (int64_t)i * cap);
  if (p >= min(pair_count[i], cap)) return;
  const int j = pairs[t];
  if (stride > 1 && topo_hash(i, j, frame) % stride != 0) return;
  if (early_exit && deg_ema[i] >= deg_high && deg_ema[j] >= deg_high) return;
  if (gabriel_pair(pos, i, j, G, L, half_l, inv_l, cell_size, cs, cc, ci) == 1 && j > i) {
    atomicAdd(&deg[i], 1);
    atomicAdd(&deg[j], 1);
  }
}

// topology.py:335-446 gabriel_update_batched, part 1: the candidate pairs of
// the batch's particles (j > i, hash-subsampled, within mult*(ri+rj)).  The
// count goes past the list like the reference's atomic counter.
__global__ void k_topo_batch_pairs(const float *__restrict__ pos, const float *__restrict__ rad,
                                   int n, int G, float L, float half_l, float inv_l, float mult,
                                   int cap, const int32_t *__restrict__ cs,
                                   const int32_t *__restrict__ cc, const int32_t *__restrict__ ci,
                                   int batch_idx, int num_batches, int q, int frame,
                                   int32_t *__restrict__ pairs, int32_t *__restrict__ pair_count,
                                   int32_t *__restrict__ truncated) {
  const int i = batch_idx + (blockIdx.x * blockDim.x + threadIdx.x) * num_batches;
  if (i >= n) return;
  const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2], ri = rad[i];
  const int cx = wrap_fast(centre_axis(px, G, L, half_l, inv_l), G);
  const int cy = wrap_fast(centre_axis(py, G, L, half_l, inv_l), G);
  const int cz = wrap_fast(centre_axis(pz, G, L, half_l, inv_l), G);
  int c = 0, tr = 0;
  for (int ax = -1; ax <= 1; ++ax)
    for (int ay = -1; ay <= 1; ++ay)
      for (int az = -1; az <= 1; ++az) {
        const int id = (wrap1(cx + ax, G) * G + wrap1(cy + ay, G)) * G + wrap1(cz + az, G);
        const int b = cs[id], e = b + cc[id];
        for (int k = b; k < e; ++k) {
          const int j = ci[k];
          if (j <= i || topo_hash(i, j, frame) % q != 0) continue;
          const float ex = gwrap(pos[3 * j] - px, L, inv_l), ey = gwrap(pos[3 * j + 1] - py, L, inv_l),
                      ez = gwrap(pos[3 * j + 2] - pz, L, inv_l);
          const float th = mult * (ri + rad[j]);
          if (d3(ex, ey, ez) <= th * th) {
            if (c < cap)
              pairs[(int64_t)i * cap + c] = j;
            else
              tr = 1;
            c += 1;
          }
        }
      }
  pair_count[i] = c;
  truncated[i] = tr;
}

// Part 2: the Euclidean Gabriel test of the stored pairs (degrees are not
// cleared: batches accumulate).
__global__ void k_topo_batch_test(const float *__restrict__ pos, int n, int G, float L,
                                  float half_l, float inv_l, float cell_size,
                                  const int32_t *__restrict__ cs, const int32_t *__restrict__ cc,
                                  const int32_t *__restrict__ ci, int batch_idx, int num_batches,
                                  const int32_t *__restrict__ pairs,
                                  const int32_t *__restrict__ pair_count, int cap,
                                  int32_t *__restrict__ deg) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t member = t / cap;
  const int p = (int)(t - member * cap);
  const int64_t i64 = batch_idx + member * num_batches;
  if (i64 >= n) return;
  const int i = (int)i64;
  if (p >= min(pair_count[i], cap)) return;
  const int j = pairs[(int64_t)i * cap + p];
  if (gabriel_pair(pos, i, j, G, L, half_l, inv_l, cell_size, cs, cc, ci) == 1) {
    atomicAdd(&deg[i], 1);
    atomicAdd(&deg[j], 1);
  }
}

// topology.py:458-504 knn_topo_degree: neighbours within (mult*ri)^2 over the
// 27 cells, counting stops at k.
__global__ void k_topo_knn(const float *__restrict__ pos, const float *__restrict__ rad, int n,
                           int G, float L, float half_l, float inv_l, float mult, int kmax,
                           const int32_t *__restrict__ cs, const int32_t *__restrict__ cc,
                           const int32_t *__restrict__ ci, int32_t *__restrict__ deg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
  const float r = mult * rad[i], th = r * r;
  const int cx = wrap_fast(centre_axis(px, G, L, half_l, inv_l), G);
  const int cy = wrap_fast(centre_axis(py, G, L, half_l, inv_l), G);
  const int cz = wrap_fast(centre_axis(pz, G, L, half_l, inv_l), G);
  int c = 0;
  for (int ax = -1; ax <= 1 && c < kmax; ++ax)
    for (int ay = -1; ay <= 1 && c < kmax; ++ay)
      for (int az = -1; az <= 1 && c < kmax; ++az) {
        const int id = (wrap1(cx + ax, G) * G + wrap1(cy + ay, G)) * G + wrap1(cz + az, G);
        const int b = cs[id], e = b + cc[id];
        for (int l = b; l < e && c < kmax; ++l) {
          const int j = ci[l];
          if (j == i) continue;
          const float ex = gwrap(pos[3 * j] - px, L, inv_l), ey = gwrap(pos[3 * j + 1] - py, L, inv_l),
                      ez = gwrap(pos[3 * j + 2] - pz, L, inv_l);
          if (d3(ex, ey, ez) <= th) c += 1;
        }
      }
  deg[i] = min(c, kmax);
}

// topology.py:266-283 update_topo_ema / :449-455 commit_topo_to_ema.
__global__ void k_topo_ema(const int32_t *__restrict__ deg, float *__restrict__ ema, int n,
                           float alpha) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ema[i] = alpha * (float)deg[i] + (1.0f - alpha) * ema[i];
}

// topology.py:287-322 update_radii_topological.
__global__ void k_topo_radii(float *__restrict__ rad, const float *__restrict__ deg_eff, int n,
                             float deg_low, float deg_high, float gain_grow, float gain_shrink,
                             float r_min, float r_max) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float d = deg_eff[i];
  float r = rad[i];
  if (d < deg_low)
    r *= (1.0f + gain_grow);
  else if (d > deg_high)
    r *= (1.0f - gain_shrink);
  rad[i] = fmaxf(r_min, fminf(r_max, r));
}

extern "C" {

int fos_topo_gabriel(const float *pos, int n, int G, float L, float cell_size,
                     const int32_t *cell_start, const int32_t *cell_count,
                     const int32_t *cell_indices, const int32_t *pairs, const int32_t *pair_count,
                     int cap, int frame, int stride, int early_exit, float deg_high,
                     const float *deg_ema, int32_t *deg, void *stream) {
  FOS_REQUIRE(n >= 0 && G >= 3 && L > 0.0f && cap >= 1 && cell_size > 0.0f && stride >= 1 &&
                  (n == 0 || (pos && cell_start && cell_count && cell_indices && pairs &&
                              pair_count && deg && (!early_exit || deg_ema))),
              "fos_topo_gabriel: bad arguments (n=%d G=%d cap=%d)", n, G, cap);
  if (n == 0) return 0;
  int rc = fos_fill_i32(deg, n, 0, stream);
  if (rc) return rc;
  hipLaunchKernelGGL(k_topo_gabriel, dim3(blocks_for((int64_t)n * cap)), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream, pos, n, G, L, (float)(0.5 * (double)L),
                     (float)(1.0 / (double)L), cell_size, cell_start, cell_count, cell_indices,
                     pairs, pair_count, cap, frame, stride, early_exit, deg_high, deg_ema, deg);
  FOS_CHECK_LAUNCH("k_topo_gabriel");
  return 0;
}

int fos_topo_gabriel_batched(const float *pos, const float *rad, int n, int G, float L,
                             float cell_size, float mult, int cap, const int32_t *cell_start,
                             const int32_t *cell_count, const int32_t *cell_indices,
                             int batch_idx, int num_batches, int q, int frame, int32_t *pairs,
                             int32_t *pair_count, int32_t *truncated, int32_t *deg,
                             void *stream) {
  FOS_REQUIRE(n >= 0 && G >= 3 && L > 0.0f && cap >= 1 && cell_size > 0.0f && q >= 1 &&
                  num_batches >= 1 && batch_idx >= 0 && batch_idx < num_batches &&
                  (n == 0 || (pos && rad && cell_start && cell_count && cell_indices && pairs &&
                              pair_count && truncated && deg)),
              "fos_topo_gabriel_batched: bad arguments (n=%d batch %d/%d)", n, batch_idx,
              num_batches);
  if (n <= batch_idx) return 0;
  hipStream_t st = (hipStream_t)stream;
  const int64_t members = (n - batch_idx + num_batches - 1) / num_batches;
  const float half_l = (float)(0.5 * (double)L), inv_l = (float)(1.0 / (double)L);
  hipLaunchKernelGGL(k_topo_batch_pairs, dim3(blocks_for(members)), dim3(FOS_BLOCK), 0, st, pos,
                     rad, n, G, L, half_l, inv_l, mult, cap, cell_start, cell_count, cell_indices,
                     batch_idx, num_batches, q, frame, pairs, pair_count, truncated);
  FOS_CHECK_LAUNCH("k_topo_batch_pairs");
  hipLaunchKernelGGL(k_topo_batch_test, dim3(blocks_for(members * cap)), dim3(FOS_BLOCK), 0, st,
                     pos, n, G, L, half_l, inv_l, cell_size, cell_start, cell_count, cell_indices,
                     batch_idx, num_batches, (const int32_t *)pairs, (const int32_t *)pair_count,
                     cap, deg);
  FOS_CHECK_LAUNCH("k_topo_batch_test");
  return 0;
}

int fos_topo_knn(const float *pos, const float *rad, int n, int G, float L, float mult, int k,
                 const int32_t *cell_start, const int32_t *cell_count,
                 const int32_t *cell_indices, int32_t *deg, void *stream) {
  FOS_REQUIRE(n >= 0 && G >= 3 && L > 0.0f && k >= 0 &&
                  (n == 0 || (pos && rad && cell_start && cell_count && cell_indices && deg)),
              "fos_topo_knn: bad arguments (n=%d G=%d)", n, G);
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_topo_knn, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, (hipStream_t)stream,
                     pos, rad, n, G, L, (float)(0.5 * (double)L), (float)(1.0 / (double)L), mult,
                     k, cell_start, cell_count, cell_indices, deg);
  FOS_CHECK_LAUNCH("k_topo_knn");
  return 0;
}

int fos_topo_ema(const int32_t *deg, float *ema, int n, float alpha, void *stream) {
  FOS_REQUIRE(n >= 0 && (n == 0 || (deg && ema)), "fos_topo_ema: bad arguments (n=%d)", n);
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_topo_ema, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, (hipStream_t)stream,
                     deg, ema, n, alpha);
  FOS_CHECK_LAUNCH("k_topo_ema");
  return 0;
}

int fos_topo_radii(float *rad, const float *deg_eff, int n, float deg_low, float deg_high,
                   float gain_grow, float gain_shrink, float r_min, float r_max, void *stream) {
  FOS_REQUIRE(n >= 0 && (n == 0 || (rad && deg_eff)), "fos_topo_radii: bad arguments (n=%d)", n);
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_topo_radii, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, (hipStream_t)stream,
                     rad, deg_eff, n, deg_low, deg_high, gain_grow, gain_shrink, r_min, r_max);
  FOS_CHECK_LAUNCH("k_topo_radii");
  return 0;
}

}  // extern "C"
