// fos_grid.hip -- periodic spatial hash as a counting sort (reference:
// grid.py:145-309, rebuild order run.py:476-504).  Cells are G^3, index
// cx*G*G + cy*G + cz.  Particles inside a cell are listed in ascending index
// (the serial scatter order), which fixes the PBD accumulation order.
#include "fos_common.h"

using namespace fos;

__device__ __forceinline__ int particle_cell(const float *__restrict__ pos, int i,
                                             int G, float L, float half_l,
                                             float inv_l, float pw[3]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) pw[d] = gwrap(pos[3 * i + d], L, inv_l);
  const int c0 = hash_axis(pw[0], G, half_l, inv_l);
  const int c1 = hash_axis(pw[1], G, half_l, inv_l);
  const int c2 = hash_axis(pw[2], G, half_l, inv_l);
  return c0 * G * G + c1 * G + c2;
}

// grid.py:202-222
__global__ void k_grid_count(const float *__restrict__ pos, int n, int G, float L,
                             float half_l, float inv_l, int32_t *__restrict__ count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float pw[3];
  const int c = particle_cell(pos, i, G, L, half_l, inv_l, pw);
  atomicAdd(&count[c], 1);
}

// grid.py:285-309: wraps pos in place and appends i to its cell.
__global__ void k_grid_scatter(float *__restrict__ pos, int n, int G, float L,
                               float half_l, float inv_l, int32_t *__restrict__ wr,
                               int32_t *__restrict__ idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float pw[3];
  const int c = particle_cell(pos, i, G, L, half_l, inv_l, pw);
  pos[3 * i] = pw[0];
  pos[3 * i + 1] = pw[1];
  pos[3 * i + 2] = pw[2];
  const int slot = atomicAdd(&wr[c], 1);
  idx[slot] = i;
}

// Restores the serial (stable) order inside each cell.  After the scatter,
// wr[c] is the end of cell c and wr[c-1] its start (cells are contiguous).
__global__ void k_grid_sort_cells(int C, const int32_t *__restrict__ wr,
                                  int32_t *__restrict__ idx) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  const int b = c ? wr[c - 1] : 0, e = wr[c];
  for (int p = b + 1; p < e; ++p) {
    const int key = idx[p];
    int q = p - 1;
    while (q >= b && idx[q] > key) {
      idx[q + 1] = idx[q];
      --q;
    }
    idx[q + 1] = key;
  }
}

// Large-system scatter: wraps pos, records each particle's cell and appends
// it to a temporary cell-ordered list (arbitrary order inside a cell).
__global__ void k_grid_scatter_tmp(float *__restrict__ pos, int n, int G, float L, float half_l,
                                   float inv_l, int32_t *__restrict__ wr,
                                   int32_t *__restrict__ cell_of, int32_t *__restrict__ tmp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float pw[3];
  const int c = particle_cell(pos, i, G, L, half_l, inv_l, pw);
  pos[3 * i] = pw[0];
  pos[3 * i + 1] = pw[1];
  pos[3 * i + 2] = pw[2];
  cell_of[i] = c;
  tmp[atomicAdd(&wr[c], 1)] = i;
}

// Places each particle at its cell's start plus the number of smaller
// indices in its cell (the stable order), one thread per particle.
__global__ void k_grid_rank(int n, const int32_t *__restrict__ cell_of,
                            const int32_t *__restrict__ start, const int32_t *__restrict__ count,
                            const int32_t *__restrict__ tmp, int32_t *__restrict__ idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cell_of[i];
  const int b = start[c], e = b + count[c];
  int r = b;
  for (int q = b; q < e; ++q) r += tmp[q] < i ? 1 : 0;
  idx[r] = i;
}

__global__ void k_copy_cells(const int32_t *__restrict__ s, int32_t *__restrict__ d,
                             int C) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < C) d[c] = s[c];
}

// grid.py:229-261 in one block: exclusive scan of up to SCAN_CELLS_MAX cell
// counts into cell_start and the scatter pointers.  Counts are staged in LDS
// with coalesced loads; each thread scans a contiguous run.
#define SCAN_CELLS_MAX 12288
__global__ void __launch_bounds__(1024)
    k_scan_cells(const int32_t *__restrict__ count, int32_t *__restrict__ start,
                 int32_t *__restrict__ wr, int C) {
  __shared__ int s_wave[16];
  __shared__ int s_c[SCAN_CELLS_MAX];
  const int t = (int)threadIdx.x, lane = t & 63, wid = t >> 6;
  for (int k = t; k < C; k += 1024) s_c[k] = count[k];
  __syncthreads();
  const int per = (C + 1023) / 1024;
  const int b = t * per, e = b + per < C ? b + per : C;
  int s = 0;
  for (int k = b; k < e; ++k) s += s_c[k];
  int x = s;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wave[wid] = x;
  __syncthreads();
  if (wid == 0) {
    int w = lane < 16 ? s_wave[lane] : 0;
    for (int o = 1; o < 16; o <<= 1) {
      const int y = __shfl_up(w, o, 64);
      if (lane >= o) w += y;
    }
    if (lane < 16) s_wave[lane] = w;
  }
  __syncthreads();
  int ex = (wid ? s_wave[wid - 1] : 0) + x - s;
  for (int k = b; k < e; ++k) {
    const int c = s_c[k];
    s_c[k] = ex;
    ex += c;
  }
  __syncthreads();
  for (int k = t; k < C; k += 1024) {
    const int v = s_c[k];
    start[k] = v;
    wr[k] = v;
  }
}

// Cell of every particle, wrapped positions written back (grid.py:202-222 and
// the wrap of grid.py:285-309), spread over many workgroups.
__global__ void k_grid_cells(float *__restrict__ pos, int n, int G, float L, float half_l,
                             float inv_l, int32_t *__restrict__ cell_of) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float pw[3];
  const int c = particle_cell(pos, i, G, L, half_l, inv_l, pw);
  pos[3 * i] = pw[0];
  pos[3 * i + 1] = pw[1];
  pos[3 * i + 2] = pw[2];
  cell_of[i] = c;
}

// Rebuild by cell ranges: workgroup w owns cells [w*cpw, (w+1)*cpw) (at
// most GR_CELLS).  It reads every particle's cell, counts its own cells and
// the particles in earlier cells (its global offset), scans its counts,
// places its particles in arbitrary order within a cell and then sorts each
// cell by particle index (the stable order of grid.py:285-309).  No
// workgroup depends on another, so one launch does count, scan and scatter.
// Up to GR_LIST particles of a range are staged in LDS; a range with more is
// placed in cell_indices directly.
#define GR_CELLS FOS_GR_CELLS
#define GR_LIST 2048

__device__ __forceinline__ void sort_run(int32_t *a, int m) {
  for (int x = 1; x < m; ++x) {
    const int v = a[x];
    int y = x - 1;
    while (y >= 0 && a[y] > v) {
      a[y + 1] = a[y];
      --y;
    }
    a[y + 1] = v;
  }
}

// Builds cells [lo, hi) of the CSR from the particles [pa, pe), whose cells
// all lie in the grid of those cells; base0 = first CSR slot of particle pa's
// grid.  All threads of the workgroup call it.
// With s4 given, s4[slot] = (position, radius) of the slot's particle (the
// cell-ordered copy the batched neighbour kernels read).
__device__ __forceinline__ void grid_range(const int32_t *__restrict__ cell_of, int pa, int pe,
                                           int base0, int lo, int hi,
                                           int32_t *__restrict__ count,
                                           int32_t *__restrict__ start,
                                           int32_t *__restrict__ idx,
                                           const float *__restrict__ pos = nullptr,
                                           const float *__restrict__ rad = nullptr,
                                           float4 *__restrict__ s4 = nullptr) {
  __shared__ int s_cnt[GR_CELLS], s_ex[GR_CELLS];
  __shared__ int s_li[GR_LIST], s_lc[GR_LIST], s_out[GR_LIST];
  __shared__ int s_w[16], s_m;
  const int t = (int)threadIdx.x, lane = t & 63, wid = t >> 6;
  const int nc = hi - lo;
  for (int k = t; k < nc; k += 1024) s_cnt[k] = 0;
  if (t == 0) s_m = 0;
  __syncthreads();
  int before = 0;
  auto take = [&](int i, int c) {
    if (c < lo) {
      ++before;
    } else if (c < hi) {
      atomicAdd(&s_cnt[c - lo], 1);
      const int k = atomicAdd(&s_m, 1);
      if (k < GR_LIST) {
        s_li[k] = i;
        s_lc[k] = c - lo;
      }
    }
  };
  // cells in 16-byte loads over the aligned middle, four per thread in
  // flight (cell_of is a 256-byte aligned context allocation)
  const int a4 = (pa + 3) & ~3, e4 = pe & ~3;
  if (a4 < e4) {
    for (int i = pa + t; i < a4; i += 1024) take(i, cell_of[i]);
    const int m4 = (e4 - a4) >> 2;
    const int4 *c4 = (const int4 *)(cell_of + a4);
    for (int b = 0; b < m4; b += 4 * 1024) {
      int4 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = b + u * 1024 + t;
        v[u] = k < m4 ? c4[k] : make_int4(hi, hi, hi, hi);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = a4 + 4 * (b + u * 1024 + t);
        take(i, v[u].x);
        take(i + 1, v[u].y);
        take(i + 2, v[u].z);
        take(i + 3, v[u].w);
      }
    }
    for (int i = e4 + t; i < pe; i += 1024) take(i, cell_of[i]);
  } else {
    for (int i = pa + t; i < pe; i += 1024) take(i, cell_of[i]);
  }
  for (int o = 32; o; o >>= 1) before += __shfl_xor(before, o, 64);
  if (lane == 0) s_w[wid] = before;
  __syncthreads();
  int base = base0;  // slots of the particles in cells before lo
#pragma unroll
  for (int w = 0; w < 16; ++w) base += s_w[w];
  const int m = s_m;
  __syncthreads();
  // exclusive scan of the range's counts, one cell per thread
  const int v = t < nc ? s_cnt[t] : 0;
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_w[wid] = x;
  __syncthreads();
  int wbase = 0;
  for (int w = 0; w < wid; ++w) wbase += s_w[w];
  if (t < nc) {
    const int ex = wbase + x - v;
    s_ex[t] = ex;
    s_cnt[t] = 0;  // becomes the placement cursor
    count[lo + t] = v;
    start[lo + t] = base + ex;
  }
  __syncthreads();
  if (m <= GR_LIST) {
    for (int k = t; k < m; k += 1024) {
      const int c = s_lc[k];
      s_out[s_ex[c] + atomicAdd(&s_cnt[c], 1)] = s_li[k];
    }
    __syncthreads();
    for (int k = t; k < nc; k += 1024) sort_run(s_out + s_ex[k], s_cnt[k]);
    __syncthreads();
    for (int k = t; k < m; k += 1024) {
      const int i = s_out[k];
      idx[base + k] = i;
      if (s4) s4[base + k] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], rad[i]);
    }
  } else {
    for (int i = pa + t; i < pe; i += 1024) {
      const int c = cell_of[i];
      if (c >= lo && c < hi) idx[base + s_ex[c - lo] + atomicAdd(&s_cnt[c - lo], 1)] = i;
    }
    __threadfence();
    __syncthreads();
    for (int k = t; k < nc; k += 1024) sort_run(idx + base + s_ex[k], s_cnt[k]);
    if (s4) {
      __threadfence();
      __syncthreads();
      for (int k = t; k < m; k += 1024) {
        const int i = idx[base + k];
        s4[base + k] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], rad[i]);
      }
    }
  }
}

// s4 (may be NULL): the slot records, written as the slots are placed
__global__ void __launch_bounds__(1024)
    k_grid_build_ranges(const int32_t *__restrict__ cell_of, int n, int C, int cpw,
                        int32_t *__restrict__ count, int32_t *__restrict__ start,
                        int32_t *__restrict__ idx, const float *__restrict__ pos,
                        const float *__restrict__ rad, float4 *__restrict__ s4) {
  const int lo = (int)blockIdx.x * cpw;
  grid_range(cell_of, 0, n, 0, lo, min(C, lo + cpw), count, start, idx, pos, rad, s4);
}

// Slot records of a built CSR: s4[k] = (position, radius) of particle idx[k].
__global__ void k_grid_records(const int32_t *__restrict__ idx, const float *__restrict__ pos,
                               const float *__restrict__ rad, int n, float4 *__restrict__ s4) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = idx[k];
  s4[k] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], rad[i]);
}

// Batches of foams: the cells of every particle in its foam's grid (global
// cell ids), skipping foams that sit this rebuild out (npass given and
// pass >= npass[b]: their positions, cells and CSR stay as they are).
__global__ void k_grid_cells_b(float *__restrict__ pos, const int32_t *__restrict__ pinst,
                               const fos_inst *__restrict__ inst, int n, float L, float half_l,
                               float inv_l, const int32_t *__restrict__ npass, int pass,
                               int32_t *__restrict__ cell_of) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = pinst[i];
  if (npass && pass >= npass[b]) return;
  float pw[3];
  const int c = particle_cell(pos, i, inst[b].G, L, half_l, inv_l, pw);
  pos[3 * i] = pw[0];
  pos[3 * i + 1] = pw[1];
  pos[3 * i + 2] = pw[2];
  cell_of[i] = inst[b].c0 + c;
}

__global__ void __launch_bounds__(1024)
    k_grid_build_ranges_b(const int32_t *__restrict__ cell_of,
                          const fos_inst *__restrict__ inst, const int *__restrict__ gstart, int B,
                          const int32_t *__restrict__ npass, int pass,
                          int32_t *__restrict__ count, int32_t *__restrict__ start,
                          int32_t *__restrict__ idx, const float *__restrict__ pos,
                          const float *__restrict__ rad, float4 *__restrict__ s4) {
  const int w = (int)blockIdx.x;
  int lo = 0, hi = B - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (gstart[mid] <= w) lo = mid; else hi = mid - 1;
  }
  const int b = lo;
  if (npass && pass >= npass[b]) return;
  const fos_inst I = inst[b];
  const int C = I.G * I.G * I.G, cpw = fos_range_cells(C);
  const int c_lo = I.c0 + (w - gstart[b]) * cpw;
  grid_range(cell_of, I.p0, I.p0 + I.n, I.p0, c_lo, min(I.c0 + C, c_lo + cpw), count, start, idx,
             pos, rad, s4);
}

// 0: multi-kernel rebuild; 1: one workgroup where it fits (else 2);
// 2: cell-range workgroups.
int g_grid_block = 1;  // fos_set_param "grid_block": 1 cell ranges, 0 the kernel chain
// cells per range workgroup of a single-foam rebuild (0: fos_range_cells)
int g_grid_cpw = 0;

static int range_cells(int C) {
  return g_grid_cpw > 0 ? (g_grid_cpw < C ? g_grid_cpw : C) : fos_range_cells(C);
}

extern "C" {

int fos_grid_count(const float *pos, int n, int G, float L, int32_t *cell_count,
                   void *stream) {
  FOS_REQUIRE(n >= 0 && G >= 3 && G <= 1290 && L > 0.0f && cell_count &&
                  (n == 0 || pos),
              "fos_grid_count: bad arguments (n=%d G=%d)", n, G);
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_grid_count, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream, pos, n, G, L, (float)(0.5 * (double)L),
                     (float)(1.0 / (double)L), cell_count);
  FOS_CHECK_LAUNCH("k_grid_count");
  return 0;
}

int fos_grid_scan(fos_ctx *ctx, const int32_t *cell_count, int32_t *cell_start,
                  int G, void *stream) {
  FOS_REQUIRE(ctx && cell_count && cell_start && G >= 3 &&
                  (int64_t)G * G * G <= ctx->max_cells,
              "fos_grid_scan: bad arguments (G=%d)", G);
  return exclusive_scan(ctx, cell_count, cell_start, G * G * G, (hipStream_t)stream);
}

int fos_grid_scatter(float *pos, int n, int G, float L, int32_t *cell_write,
                     int32_t *cell_indices, void *stream) {
  FOS_REQUIRE(n >= 0 && G >= 3 && G <= 1290 && L > 0.0f && cell_write &&
                  (n == 0 || (pos && cell_indices)),
              "fos_grid_scatter: bad arguments (n=%d G=%d)", n, G);
  if (n == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_grid_scatter, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, st, pos,
                     n, G, L, (float)(0.5 * (double)L), (float)(1.0 / (double)L),
                     cell_write, cell_indices);
  FOS_CHECK_LAUNCH("k_grid_scatter");
  const int C = G * G * G;
  hipLaunchKernelGGL(k_grid_sort_cells, dim3(blocks_for(C)), dim3(FOS_BLOCK), 0, st, C,
                     cell_write, cell_indices);
  FOS_CHECK_LAUNCH("k_grid_sort_cells");
  return 0;
}

static int grid_records(const float *pos, const float *rad, int n, const int32_t *cell_indices,
                        float *cell_pr, hipStream_t st) {
  if (!cell_pr || n == 0) return 0;
  hipLaunchKernelGGL(k_grid_records, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, st, cell_indices,
                     pos, rad, n, (float4 *)cell_pr);
  FOS_CHECK_LAUNCH("k_grid_records");
  return 0;
}

int fos_grid_build_from_cells(fos_ctx *ctx, const float *pos, const float *rad, int n, int G,
                              int32_t *cell_count, int32_t *cell_start, int32_t *cell_indices,
                              float *cell_pr, void *stream) {
  FOS_REQUIRE(ctx && G >= 3 && (int64_t)G * G * G <= ctx->max_cells && n >= 0 &&
                  n <= ctx->max_n && cell_count && cell_start && (n == 0 || cell_indices) &&
                  (!cell_pr || n == 0 || (pos && rad)),
              "fos_grid_build_from_cells: bad arguments (n=%d G=%d)", n, G);
  const int C = G * G * G;
  hipStream_t st = (hipStream_t)stream;
  const int cpw = range_cells(C);
  const int wgs = (C + cpw - 1) / cpw;
  hipLaunchKernelGGL(k_grid_build_ranges, dim3((unsigned)wgs), dim3(1024), 0, st,
                     (const int32_t *)ctx->cell_of, n, C, cpw, cell_count, cell_start,
                     cell_indices, cell_pr ? pos : nullptr, cell_pr ? rad : nullptr,
                     (float4 *)cell_pr);
  FOS_CHECK_LAUNCH("k_grid_build_ranges");
  return 0;
}

int fos_grid_build_pr(fos_ctx *ctx, float *pos, const float *rad, int n, int G, float L,
                      int32_t *cell_count, int32_t *cell_start, int32_t *cell_indices,
                      float *cell_pr, void *stream) {
  FOS_REQUIRE(!cell_pr || n == 0 || rad, "fos_grid_build_pr: NULL rad");
  const int C = G * G * G;
  if (cell_pr && n > 0 && g_grid_block == 1 && ctx && G >= 3 &&
      (int64_t)C <= ctx->max_cells && n <= ctx->max_n && pos && cell_indices && cell_count &&
      cell_start) {
    // the range form writes the records as it places the slots
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_grid_cells, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, st, pos, n, G, L,
                       (float)(0.5 * (double)L), (float)(1.0 / (double)L), ctx->cell_of);
    FOS_CHECK_LAUNCH("k_grid_cells");
    return fos_grid_build_from_cells(ctx, pos, rad, n, G, cell_count, cell_start, cell_indices,
                                     cell_pr, stream);
  }
  int rc = fos_grid_build(ctx, pos, n, G, L, cell_count, cell_start, cell_indices, stream);
  if (rc) return rc;
  return grid_records(pos, rad, n, cell_indices, cell_pr, (hipStream_t)stream);
}

int fos_grid_build(fos_ctx *ctx, float *pos, int n, int G, float L,
                   int32_t *cell_count, int32_t *cell_start,
                   int32_t *cell_indices, void *stream) {
  FOS_REQUIRE(ctx && G >= 3 && (int64_t)G * G * G <= ctx->max_cells && n >= 0 &&
                  n <= ctx->max_n && cell_count && cell_start,
              "fos_grid_build: bad arguments (n=%d G=%d)", n, G);
  const int C = G * G * G;
  hipStream_t st = (hipStream_t)stream;
  if (g_grid_block) {
    FOS_REQUIRE(n == 0 || (cell_indices && pos), "fos_grid_build: NULL pos or cell_indices");
    if (n > 0) {
      hipLaunchKernelGGL(k_grid_cells, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, st, pos, n, G,
                         L, (float)(0.5 * (double)L), (float)(1.0 / (double)L), ctx->cell_of);
      FOS_CHECK_LAUNCH("k_grid_cells");
    }
    const int cpw = range_cells(C);
    const int wgs = (C + cpw - 1) / cpw;
    hipLaunchKernelGGL(k_grid_build_ranges, dim3((unsigned)wgs), dim3(1024), 0, st,
                       (const int32_t *)ctx->cell_of, n, C, cpw, cell_count, cell_start,
                       cell_indices, (const float *)nullptr, (const float *)nullptr,
                       (float4 *)nullptr);
    FOS_CHECK_LAUNCH("k_grid_build_ranges");
    return 0;
  }
  int rc = fos_fill_i32(cell_count, C, 0, stream);
  if (rc) return rc;
  rc = fos_grid_count(pos, n, G, L, cell_count, stream);
  if (rc) return rc;
  if (C <= SCAN_CELLS_MAX) {
    hipLaunchKernelGGL(k_scan_cells, dim3(1), dim3(1024), 0, st, cell_count, cell_start,
                       ctx->cell_write, C);
    FOS_CHECK_LAUNCH("k_scan_cells");
  } else {
    rc = exclusive_scan(ctx, cell_count, cell_start, C, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_copy_cells, dim3(blocks_for(C)), dim3(FOS_BLOCK), 0, st,
                       cell_start, ctx->cell_write, C);
    FOS_CHECK_LAUNCH("k_copy_cells");
  }
  if (n == 0) return 0;
  FOS_REQUIRE(pos && cell_indices, "fos_grid_build: NULL pos or cell_indices");
  const float half_l = (float)(0.5 * (double)L), inv_l = (float)(1.0 / (double)L);
  hipLaunchKernelGGL(k_grid_scatter_tmp, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, st, pos, n, G,
                     L, half_l, inv_l, ctx->cell_write, ctx->cell_of, ctx->grid_tmp);
  FOS_CHECK_LAUNCH("k_grid_scatter_tmp");
  hipLaunchKernelGGL(k_grid_rank, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, st, n,
                     (const int32_t *)ctx->cell_of, (const int32_t *)cell_start,
                     (const int32_t *)cell_count, (const int32_t *)ctx->grid_tmp, cell_indices);
  FOS_CHECK_LAUNCH("k_grid_rank");
  return 0;
}

}  // extern "C"

extern "C" {

int fos_batch_grid_build(fos_batch *bt, fos_ctx *ctx, float *pos, float L, int32_t *cell_count,
                         int32_t *cell_start, int32_t *cell_indices, const int32_t *passes,
                         int pass, const float *rad, float *cell_pr, void *stream) {
  FOS_REQUIRE(bt && ctx && pos && cell_count && cell_start && cell_indices && L > 0.0f &&
                  bt->n_total <= ctx->max_n && (!cell_pr || rad),
              "fos_batch_grid_build: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const int n = bt->n_total;
  hipLaunchKernelGGL(k_grid_cells_b, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, st, pos, bt->pinst,
                     bt->inst, n, L, (float)(0.5 * (double)L), (float)(1.0 / (double)L), passes,
                     pass, ctx->cell_of);
  FOS_CHECK_LAUNCH("k_grid_cells_b");
  hipLaunchKernelGGL(k_grid_build_ranges_b, dim3((unsigned)bt->gstart_h[bt->B]), dim3(1024), 0,
                     st, (const int32_t *)ctx->cell_of, bt->inst, bt->gstart, bt->B, passes, pass,
                     cell_count, cell_start, cell_indices, (const float *)pos, rad,
                     (float4 *)cell_pr);
  FOS_CHECK_LAUNCH("k_grid_build_ranges_b");
  return 0;
}

}  // extern "C"
