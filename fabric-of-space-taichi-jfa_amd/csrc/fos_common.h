// fos_common.h -- device helpers and launch plumbing shared by the HIP
// translation units of libfos.so.
//
// Numerics contract: every fp32 expression below is evaluated in the order
// the reference writes it, with contraction disabled (the library is built
// with -ffp-contract=off) and HIP's correctly rounded fp32 division/sqrt.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fos.h"

#define FOS_MAX_NEIGHBORS 32
#define FOS_BLOCK 256

namespace fos {

// Sets the thread-local error text returned by fos_last_error().
void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);

#define FOS_CHECK_LAUNCH(name)                                   \
  do {                                                           \
    hipError_t _e = hipGetLastError();                           \
    if (_e != hipSuccess) return ::fos::hip_fail(_e, name);      \
  } while (0)

#define FOS_HIP(call)                                            \
  do {                                                           \
    hipError_t _e = (call);                                      \
    if (_e != hipSuccess) return ::fos::hip_fail(_e, #call);     \
  } while (0)

#define FOS_REQUIRE(cond, ...)                                   \
  do {                                                           \
    if (!(cond)) {                                               \
      ::fos::set_error(__VA_ARGS__);                             \
      return FOS_EINVAL;                                         \
    }                                                            \
  } while (0)

static inline unsigned blocks_for(int64_t n, int bs = FOS_BLOCK) {
  return (unsigned)((n + bs - 1) / bs);
}

// jfa.py:247-251 / grid.py:91-95: ((i % r) + r) % r
__device__ __forceinline__ int wrapi(int i, int r) { return ((i % r) + r) % r; }

// Wrap for |off| <= r: one conditional add/sub equals the double modulo.
__device__ __forceinline__ int wrap1(int i, int r) {
  return i < 0 ? i + r : (i >= r ? i - r : i);
}

// jfa.py:303-308 wrapP for the power distance: d - L*round(d/L).  With
// L == 1 the division and product are exact and round/rint differ only at
// |d| == 0.5, where both candidates have the same square, so the squared
// distance is bit-identical with rint (one instruction).
template <bool UNIT_L>
__device__ __forceinline__ float jwrap(float d, float L) {
  if (UNIT_L) return d - rintf(d);
  return d - L * roundf(d / L);
}

// jfa.py:285-287 voxel_to_world
__device__ __forceinline__ float vcenter(int i, float voxel, float half_l) {
  return ((float)i + 0.5f) * voxel - half_l;
}

// jfa.py:451-477 power_distance_sq with a precomputed seed (x, y, z,
// (beta*r)*(beta*r)); the dot product sums left to right.
template <bool UNIT_L>
__device__ __forceinline__ float power_dist(float cx, float cy, float cz,
                                            float4 s, float L) {
  float dx = jwrap<UNIT_L>(cx - s.x, L);
  float dy = jwrap<UNIT_L>(cy - s.y, L);
  float dz = jwrap<UNIT_L>(cz - s.z, L);
  float d2 = __fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy));
  d2 = __fadd_rn(d2, __fmul_rn(dz, dz));
  return __fsub_rn(d2, s.w);
}

// grid.py:30-48 wrapP: p - L*floor(p*INV_L + 0.5)
__device__ __forceinline__ float gwrap(float p, float L, float inv_l) {
  return p - L * floorf(p * inv_l + 0.5f);
}

// wrapi with the in-range case first (no integer division there).
__device__ __forceinline__ int wrap_fast(int i, int r) {
  return (i >= 0 && i < r) ? i : wrapi(i, r);
}

// grid.py:202-219 cell of a wrapped coordinate, with the 1e-7 hash epsilon.
__device__ __forceinline__ int hash_axis(float pw, int G, float half_l,
                                         float inv_l) {
  float q = ((pw + half_l) + 1e-7f) * inv_l;
  return wrap_fast((int)floorf(q * (float)G), G);
}

// dynamics.py:101-103 stencil centre: no epsilon, truncating cast, no wrap.
__device__ __forceinline__ int centre_axis(float p, int G, float L,
                                           float half_l, float inv_l) {
  float pw = gwrap(p, L, inv_l);
  float q = (pw + half_l) * inv_l;
  return (int)(q * (float)G);
}

// Philox4x32-10 (Salmon et al., SC'11).
__device__ __forceinline__ uint4 philox(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
    uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
    uint4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

__device__ __forceinline__ float u01(uint32_t b) {
  return (float)(b >> 8) * 5.9604644775390625e-8f;
}
__device__ __forceinline__ float u01_open(uint32_t b) {
  return (float)((b >> 8) + 1u) * 5.9604644775390625e-8f;
}

// Per-pass change counters are striped over FOS_STRIPES cache lines (block id
// mod FOS_STRIPES): one hot address would serialise every block's atomic.
#define FOS_STRIPES 64
#define FOS_STRIPE_INTS 16
#define FOS_PASS_SLOTS 64
#define FOS_STRIPE_PASS_INTS (FOS_STRIPES * FOS_STRIPE_INTS)

// Wave-wide total of pass q's striped counter (every lane gets it).
__device__ __forceinline__ int stripes_total(const int *stripes, int q) {
  const int lane = (int)(threadIdx.x & 63);
  int v = stripes[(q * FOS_STRIPES + lane) * FOS_STRIPE_INTS];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block-wide sum of an int into one atomic per block.
__device__ __forceinline__ void block_add(int v, int *dst) {
  __shared__ int s_acc;
  if (threadIdx.x == 0 && threadIdx.y == 0 && threadIdx.z == 0) s_acc = 0;
  __syncthreads();
  // wave-level reduction first
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((((threadIdx.z * blockDim.y + threadIdx.y) * blockDim.x + threadIdx.x) & 63) == 0 && v)
    atomicAdd(&s_acc, v);
  __syncthreads();
  if (threadIdx.x == 0 && threadIdx.y == 0 && threadIdx.z == 0 && s_acc)
    atomicAdd(dst, s_acc);
}

}  // namespace fos

// Scratch owned by a context (fos_ctx_create); sized once, never per call.
// scheduler statistics (fos_sched_stats): FOS_SCHED_BLOCKS block partials of
// 8 doubles, the 8 first-stage totals, the 8 second-stage totals, then the 8
// doubles of fos_jfa_schedule's result
#define FOS_SCHED_BLOCKS 256
#define FOS_SCHED_DOUBLES ((FOS_SCHED_BLOCKS + 3) * 8)

struct fos_ctx {
  int device;
  int max_n;
  int max_res;
  int max_cells;
  int fin_burst;  // round pairs the last finalize needed (0: none yet); sizes the first burst
  int stripes_clear;  // 1: fos_jfa_init has queued the clear of pass slots 0..FOS_PASS_SLOTS-2
  // seed table (x, y, z, (beta*r)^2)
  float4 *seeds;
  // each voxel's winning power distance in the last two JFA passes [max_res^3]
  // (NULL: the passes evaluate every incumbent)
  float *dist[2];
  // striped per-pass changed-voxel counters [FOS_PASS_SLOTS][FOS_STRIPES][16]
  int *stripes;
  // face-pair row tables: capacity tab_cap entries per particle row
  int tab_cap;
  int *tab_b;
  unsigned *tab_k;
  int *tab_c;
  int *flags;  // [0] row-table overflow, [3] sticky finalize failure,
               // [4..5] warm-start gate counters (zero between calls),
               // [16..63] per-pair change flags of the finalize fixed point
  // finalize scratch
  float *score;           // [max_n*32]
  unsigned char *cls;     // [max_n*32]
  unsigned char *dec;     // [max_n*32]
  int *in_count;          // [max_n+1]
  int *in_start;          // [max_n+1]
  int *in_fill;           // [max_n]
  int *in_src;            // [max_n*32]
  int *in_sorted;         // [max_n*32]
  int *acc_in;            // [max_n]
  // scan scratch
  int *scan_blocks;       // block partials
  int scan_blocks_cap;
  // grid scratch
  int *cell_write;        // [max_cells]
  int *cell_of;           // [max_n]
  int *grid_tmp;          // [max_n] cell-ordered list before the stable placement
  float *pos_tmp;         // [max_n*3]
  float *rad_tmp;         // [max_n] pressure ping-pong
  // pressure iterations of one launch: slot[k][i] = (tag << 32) | radius
  // bits of particle i after iteration k, [FOS_FLOW_MAX_ITERS][flow_n]
  unsigned long long *flow;
  int flow_n;             // particles the slots hold: min(max_n, FOS_FLOW_MAX_N)
  unsigned flow_tag;      // tag of the last such launch (0: none; slots start at 0)
  int flow_blocks;        // co-resident blocks of the 16-lane form (0: not asked yet, -1: none)
  int flow_blocks_lane;   // the same for the one-lane form
  // a face step queued without its host read (fos::face_step_defer): the
  // flags' copy event, and the first burst's round count
  void *spec_event;
  int spec_pending;
  int spec_first;
  double *sched;          // [FOS_SCHED_DOUBLES] scheduler statistics: block partials, totals
  float *fscratch;        // [16] small device floats (reductions)
  int *iscratch;          // [16] small device ints
  // host pinned mirror for small reads
  int64_t *host_i;
  double *host_f;
};

// One foam of a batch (device and host descriptor).  Particles, labels and
// hash cells of the batch's foams are concatenated; particle and label ids
// are batch-global, so kernels that only follow ids (face rows, finalize,
// pressure partners) need no instance data.
struct fos_inst {
  int n, p0;           // particles [p0, p0 + n)
  int res, passes;     // JFA grid edge and scheduled passes
  long long v0;        // labels [v0, v0 + res^3)
  float voxel;         // L / res
  int G, c0;           // hash grid edge, cells [c0, c0 + G^3)
  float r_min, r_max;  // radius bounds
  float f_low, f_high; // FSC band
  unsigned seed;       // noise seed
  int tile;            // JFA pass tile: 0 rows (4 x 64 z), 1 flat (256 voxels of a plane)
};

#define FOS_BATCH_PASSES 16  // pass counter slots per foam of a batch
#define FOS_FLOW_MAX_ITERS 32  // pressure iterations one launch can run
#define FOS_FLOW_MAX_N 131072  // particles its slots hold
#define FOS_FLAG_FLOW 8        // ctx->flags slot: a pressure dependency wait timed out
namespace fos {
// The arguments of fos_jfa_faces_controller.
struct FaceStep {
  int res, n;
  const int32_t *filtered;
  float *rad_target;
  const float *rad;
  float voxel;
  int32_t *face_ids, *face_counts;
  float *prev_scores;
  int32_t *neighbor_list, *neighbor_count, *used_count, *fsc;
  float *fsc_ema, alpha;
  int frame, warmstart_frames;
  double fsc_zero_rate;
  int f_low, f_high;
  float growth_pct, r_min, r_max, deadband;
  float *gate;
};
// fos_jfa_faces_controller without its host read: the finalize's first
// burst, then materialize and the gate held on the device unless the flags
// show no row-table overflow and a fixed point.  face_step_finish (before
// anything reads the step's results) waits for the flags and, when they
// held the step back, finishes it on the host the way the synchronous call
// would; the same bits either way.
int face_step_defer(fos_ctx *ctx, const FaceStep &a, hipStream_t st);
int face_step_finish(fos_ctx *ctx, const FaceStep &a, hipStream_t st);
// pressure iterations in one launch (fos_set_param "pressure_flow")
int pressure_flow_set(int on);
int pressure_flow_get();
// lanes per particle of the slot-record PBD pass (fos_set_param "pbd_lanes")
int pbd_lanes_set(int lanes);
int pbd_lanes_get();
}  // namespace fos

// Tile -> block mapping that keeps each XCD on a contiguous range of tiles
// (workgroups are dispatched round-robin over the 8 XCDs, so consecutive
// block ids land on different L2s).
#define FOS_XCDS 8
__device__ __forceinline__ int xcd_tile(int bid, int per_xcd) {
  return (bid % FOS_XCDS) * per_xcd + bid / FOS_XCDS;
}

// Face-emission tile of a batched foam: EMIT_B_TX x-planes x (256 /
// EMIT_B_ZW) y-rows x EMIT_B_ZW z-voxels per workgroup.
#define EMIT_B_TX 8
#define EMIT_B_ZW 32

// Cells per grid-rebuild workgroup for a hash grid of C cells: about 192
// workgroups per grid, at least 64 and at most 1024 cells each.
// most cells one range workgroup of the grid rebuild takes (one per thread)
#define FOS_GR_CELLS 1024
__host__ __device__ inline int fos_range_cells(int C) {
  int cpw = (C + 191) / 192;
  return cpw < 64 ? 64 : (cpw > 1024 ? 1024 : cpw);
}

// A batch of foams: the descriptors and the per-foam work tables.
struct fos_batch {
  int B, n_total, c_total, max_res, max_passes, max_G;
  long long v_total;
  fos_inst *inst;          // device [B]
  fos_inst *inst_h;        // host [B]
  int *tstart;             // device [B+1] first JFA pass tile (row or flat, per foam)
  int *mstart;             // device [B+1] first majority tile (1 x 4 x 64 voxels)
  int *estart;             // device [B+1] first face-emission tile (EMIT_B_* shape)
  int *gstart;             // device [B+1] first grid-rebuild workgroup
  int *tstart_h, *mstart_h, *estart_h, *gstart_h;  // host copies
  int *pinst;              // device [n_total] foam of each particle
  float *pvox;             // device [n_total] voxel edge of each particle's foam
  int *stripes;            // device [B][FOS_BATCH_PASSES][FOS_STRIPE_PASS_INTS]
  float *fred;             // device [B][8] per-foam float reductions
  int *ired;               // device [B][8] per-foam int reductions
};

namespace fos {
int exclusive_scan(fos_ctx *ctx, const int *in, int *out, int n,
                   hipStream_t st);
}
