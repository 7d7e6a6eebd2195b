// fos_ctx.hip -- context lifetime, error reporting, fills and the
// device-wide exclusive scan used by the hash grid and the face transpose.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "fos_common.h"

namespace fos {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int hip_fail(hipError_t e, const char *what) {
  set_error("%s: %s", what, hipGetErrorString(e));
  return -(1000 + (int)e);
}

// ---- scan: 1024 threads x 4 items per block -----------------------------
constexpr int SCAN_T = 1024;
constexpr int SCAN_I = 4;
constexpr int SCAN_TILE = SCAN_T * SCAN_I;

__device__ __forceinline__ int block_exclusive_scan(int v, int *total) {
  __shared__ int s_wave[SCAN_T / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wave[wid] = x;
  __syncthreads();
  if (wid == 0) {
    int w = lane < SCAN_T / 64 ? s_wave[lane] : 0;
    for (int o = 1; o < 64; o <<= 1) {
      int y = __shfl_up(w, o, 64);
      if (lane >= o) w += y;
    }
    if (lane < SCAN_T / 64) s_wave[lane] = w;
  }
  __syncthreads();
  int base = wid ? s_wave[wid - 1] : 0;
  if (total) *total = s_wave[SCAN_T / 64 - 1];
  __syncthreads();
  return base + x - v;
}

__global__ void __launch_bounds__(SCAN_T) k_scan_tiles(const int *__restrict__ in,
                                                       int *__restrict__ out, int n,
                                                       int *__restrict__ block_sums) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_I;
  int v[SCAN_I];
  int s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_I; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  int total;
  int ex = block_exclusive_scan(s, &total);
#pragma unroll
  for (int k = 0; k < SCAN_I; ++k) {
    if (base + k < n) out[base + k] = ex;
    ex += v[k];
  }
  if (block_sums && threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(SCAN_T) k_scan_add(int *__restrict__ out, int n,
                                                     const int *__restrict__ offs) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_I;
  const int add = offs[blockIdx.x];
#pragma unroll
  for (int k = 0; k < SCAN_I; ++k)
    if (base + k < n) out[base + k] += add;
}

// One block scans up to SCAN_BLOCK_MAX elements staged in LDS (coalesced
// loads and stores; each thread scans a contiguous run).
constexpr int SCAN_BLOCK_MAX = 12288;
__global__ void __launch_bounds__(SCAN_T) k_scan_block(const int *__restrict__ in,
                                                       int *__restrict__ out, int n) {
  __shared__ int s_v[SCAN_BLOCK_MAX];
  const int t = (int)threadIdx.x;
  for (int k = t; k < n; k += SCAN_T) s_v[k] = in[k];
  __syncthreads();
  const int per = (n + SCAN_T - 1) / SCAN_T;
  const int b = t * per, e = b + per < n ? b + per : n;
  int s = 0;
  for (int k = b; k < e; ++k) s += s_v[k];
  int ex = block_exclusive_scan(s, nullptr);
  for (int k = b; k < e; ++k) {
    const int c = s_v[k];
    s_v[k] = ex;
    ex += c;
  }
  __syncthreads();
  for (int k = t; k < n; k += SCAN_T) out[k] = s_v[k];
}

int exclusive_scan(fos_ctx *ctx, const int *in, int *out, int n,
                   hipStream_t st) {
  if (n <= 0) return 0;
  const int tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (n <= SCAN_BLOCK_MAX) {
    hipLaunchKernelGGL(k_scan_block, dim3(1), dim3(SCAN_T), 0, st, in, out, n);
    FOS_CHECK_LAUNCH("scan block");
    return 0;
  }
  if (tiles == 1) {
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_T), 0, st, in, out, n,
                       (int *)nullptr);
    FOS_CHECK_LAUNCH("scan");
    return 0;
  }
  FOS_REQUIRE(tiles <= SCAN_TILE && tiles <= ctx->scan_blocks_cap,
              "exclusive_scan: %d elements exceed the scan capacity", n);
  int *sums = ctx->scan_blocks;
  int *offs = ctx->scan_blocks + SCAN_TILE;
  hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(SCAN_T), 0, st, in, out, n, sums);
  FOS_CHECK_LAUNCH("scan tiles");
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_T), 0, st, sums, offs, tiles,
                     (int *)nullptr);
  FOS_CHECK_LAUNCH("scan sums");
  hipLaunchKernelGGL(k_scan_add, dim3(tiles), dim3(SCAN_T), 0, st, out, n, offs);
  FOS_CHECK_LAUNCH("scan add");
  return 0;
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_f32(float *p, int64_t n, float v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_u32(unsigned *p, int64_t n, unsigned v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

static unsigned fill_blocks(int64_t n) {
  int64_t b = (n + FOS_BLOCK - 1) / FOS_BLOCK;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return (unsigned)b;
}

int fill_u32(unsigned *p, int64_t n, unsigned v, hipStream_t st) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_fill_u32, dim3(fill_blocks(n)), dim3(FOS_BLOCK), 0, st, p, n, v);
  FOS_CHECK_LAUNCH("fill_u32");
  return 0;
}

}  // namespace fos

using namespace fos;

extern "C" {

int fos_abi_version(void) { return FOS_ABI_VERSION; }

const char *fos_last_error(void) { return g_err; }

int fos_fill_i32(int32_t *dst, int64_t count, int32_t value, void *stream) {
  FOS_REQUIRE(count >= 0 && (count == 0 || dst), "fos_fill_i32: bad arguments");
  if (count == 0) return 0;
  hipLaunchKernelGGL(k_fill_i32, dim3(fill_blocks(count)), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream, dst, count, value);
  FOS_CHECK_LAUNCH("fos_fill_i32");
  return 0;
}

int fos_fill_f32(float *dst, int64_t count, float value, void *stream) {
  FOS_REQUIRE(count >= 0 && (count == 0 || dst), "fos_fill_f32: bad arguments");
  if (count == 0) return 0;
  hipLaunchKernelGGL(k_fill_f32, dim3(fill_blocks(count)), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream, dst, count, value);
  FOS_CHECK_LAUNCH("fos_fill_f32");
  return 0;
}

int fos_ctx_destroy(fos_ctx *ctx) {
  if (!ctx) return 0;
  void *ptrs[] = {ctx->seeds,    ctx->stripes,  ctx->tab_b,    ctx->tab_k,     ctx->tab_c,
                  ctx->flags,    ctx->score,    ctx->cls,       ctx->dec,
                  ctx->in_count, ctx->in_start, ctx->in_fill,   ctx->in_src,
                  ctx->in_sorted,
                  ctx->acc_in,   ctx->scan_blocks, ctx->cell_write, ctx->cell_of,
                  ctx->pos_tmp,  ctx->rad_tmp,  ctx->fscratch, ctx->iscratch,
                  ctx->grid_tmp, ctx->dist[0], ctx->dist[1],
                  ctx->flow,     ctx->sched};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (ctx->spec_event) (void)hipEventDestroy((hipEvent_t)ctx->spec_event);
  if (ctx->host_i) (void)hipHostFree(ctx->host_i);
  if (ctx->host_f) (void)hipHostFree(ctx->host_f);
  delete ctx;
  return 0;
}

}  // extern "C"

namespace fos {
// Allocates (or re-allocates) the face row tables at capacity cap per row,
// all entries empty.
int alloc_tables(fos_ctx *ctx, int cap) {
  if (ctx->tab_b) (void)hipFree(ctx->tab_b);
  if (ctx->tab_k) (void)hipFree(ctx->tab_k);
  if (ctx->tab_c) (void)hipFree(ctx->tab_c);
  ctx->tab_b = nullptr;
  ctx->tab_k = nullptr;
  ctx->tab_c = nullptr;
  const int64_t m = (int64_t)ctx->max_n * cap;
  FOS_HIP(hipMalloc(&ctx->tab_b, m * sizeof(int)));
  FOS_HIP(hipMalloc(&ctx->tab_k, m * sizeof(unsigned)));
  FOS_HIP(hipMalloc(&ctx->tab_c, m * sizeof(int)));
  ctx->tab_cap = cap;
  int rc = fos_fill_i32(ctx->tab_b, m, -1, nullptr);
  if (rc) return rc;
  rc = fill_u32(ctx->tab_k, m, 0xffffffffu, nullptr);
  if (rc) return rc;
  rc = fos_fill_i32(ctx->tab_c, m, 0, nullptr);
  if (rc) return rc;
  FOS_HIP(hipDeviceSynchronize());
  return 0;
}
}  // namespace fos

extern "C" {

int fos_ctx_create(fos_ctx **out, int max_n, int max_res, int max_cells) {
  FOS_REQUIRE(out, "fos_ctx_create: out is NULL");
  FOS_REQUIRE(max_n >= 1 && max_res >= 1 && max_res <= 894 && max_cells >= 27,
              "fos_ctx_create: need max_n>=1, 1<=max_res<=894, max_cells>=27 "
              "(got %d, %d, %d)", max_n, max_res, max_cells);
  *out = nullptr;
  fos_ctx *c = new fos_ctx();
  memset(c, 0, sizeof(*c));
  FOS_HIP(hipGetDevice(&c->device));
  c->max_n = max_n;
  c->max_res = max_res;
  c->max_cells = max_cells;
  const int64_t nr = (int64_t)max_n * FOS_MAX_NEIGHBORS;
  int rc = 0;
#define ALLOC(ptr, bytes)                                  \
  do {                                                     \
    hipError_t e = hipMalloc((void **)&(ptr), (bytes));    \
    if (e != hipSuccess) {                                 \
      rc = hip_fail(e, "fos_ctx_create hipMalloc");        \
      fos_ctx_destroy(c);                                  \
      return rc;                                           \
    }                                                      \
  } while (0)
  ALLOC(c->seeds, (size_t)max_n * sizeof(float4));
  {
    // per-voxel power distances carried between JFA passes, for grids up to
    // 256^3 (there the carry was measured faster; beyond, the extra label-
    // side traffic costs more than the evaluation it saves)
    if (max_res <= 256) {
      const size_t vox = (size_t)max_res * max_res * max_res;
      ALLOC(c->dist[0], vox * sizeof(float));
      ALLOC(c->dist[1], vox * sizeof(float));
    }
  }
  ALLOC(c->stripes, (size_t)FOS_PASS_SLOTS * FOS_STRIPE_PASS_INTS * sizeof(int));
  ALLOC(c->flags, 64 * sizeof(int));
  ALLOC(c->score, nr * sizeof(float));
  ALLOC(c->cls, nr);
  ALLOC(c->dec, nr);
  ALLOC(c->in_count, ((size_t)max_n + 1) * sizeof(int));
  ALLOC(c->in_start, ((size_t)max_n + 1) * sizeof(int));
  ALLOC(c->in_fill, (size_t)max_n * sizeof(int));
  ALLOC(c->in_src, nr * sizeof(int));
  ALLOC(c->in_sorted, nr * sizeof(int));
  ALLOC(c->acc_in, (size_t)max_n * sizeof(int));
  int64_t biggest = max_cells > max_n + 1 ? max_cells : max_n + 1;
  c->scan_blocks_cap = (int)((biggest + SCAN_TILE - 1) / SCAN_TILE) + 1;
  ALLOC(c->scan_blocks, 2 * (size_t)SCAN_TILE * sizeof(int));
  ALLOC(c->cell_write, (size_t)max_cells * sizeof(int));
  ALLOC(c->cell_of, (size_t)max_n * sizeof(int));
  ALLOC(c->grid_tmp, (size_t)max_n * sizeof(int));
  ALLOC(c->pos_tmp, (size_t)max_n * 3 * sizeof(float));
  ALLOC(c->rad_tmp, (size_t)max_n * sizeof(float));
  c->flow_n = max_n < FOS_FLOW_MAX_N ? max_n : FOS_FLOW_MAX_N;
  ALLOC(c->flow, (size_t)FOS_FLOW_MAX_ITERS * c->flow_n * sizeof(unsigned long long));
  ALLOC(c->fscratch, 64 * sizeof(float));
  ALLOC(c->iscratch, 64 * sizeof(int));
  ALLOC(c->sched, (size_t)FOS_SCHED_DOUBLES * sizeof(double));
#undef ALLOC
  hipError_t e = hipHostMalloc((void **)&c->host_i, 64 * sizeof(int64_t));
  if (e == hipSuccess) e = hipHostMalloc((void **)&c->host_f, 64 * sizeof(double));
  if (e != hipSuccess) {
    rc = hip_fail(e, "fos_ctx_create hipHostMalloc");
    fos_ctx_destroy(c);
    return rc;
  }
  rc = alloc_tables(c, 64);
  if (rc) {
    fos_ctx_destroy(c);
    return rc;
  }
  rc = fos_fill_i32(c->flags, 64, 0, nullptr);
  if (!rc)
    rc = fos_fill_i32((int32_t *)c->flow, (int64_t)2 * FOS_FLOW_MAX_ITERS * c->flow_n, 0, nullptr);
  if (rc) {
    fos_ctx_destroy(c);
    return rc;
  }
  e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    rc = hip_fail(e, "fos_ctx_create sync");
    fos_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return 0;
}

}  // extern "C"
