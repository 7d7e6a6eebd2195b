// fos_jfa.hip -- power-diagram Jump Flood and face-sharing-count extraction
// (reference: jfa.py).  Labels are int32 owner ids in one dense res^3 grid,
// index x*res*res + y*res + z (z fastest, as Taichi's ndrange).
#include <string.h>

#include "fos_common.h"

namespace fos {
int fill_u32(unsigned *p, int64_t n, unsigned v, hipStream_t st);
int alloc_tables(fos_ctx *ctx, int cap);
// hold_a / hold_b (may be NULL): device words; when either is set the gate
// and the targets are left alone (a deferred face step is not final yet).
int gate_targets(const float *rad, float *rad_target, const int32_t *fsc, const float *fsc_ema,
                 int n, int frame, int warmstart, double zero_rate, const int *zero_cnt,
                 int f_low, int f_high, float growth_pct, float r_min, float r_max,
                 float deadband, float *gate, hipStream_t st, const int *hold_a = nullptr,
                 const int *hold_b = nullptr);
}  // namespace fos

using namespace fos;

// ---------------------------------------------------------------------------
// Seed table and seeding (jfa.py:485-520)
// ---------------------------------------------------------------------------

__global__ void k_seed_table(const float *__restrict__ pos,
                             const float *__restrict__ rad, int n, float beta,
                             float4 *__restrict__ seeds) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float b = beta * rad[i];
  seeds[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], b * b);
}

// Serial semantics of the seeding loop: the owner of a voxel is the
// lexicographic minimum of (power distance to the voxel centre, index) over
// the particles seeded there.  A CAS loop reaches that minimum under any
// interleaving because the order is total.
// The seed table and the seeding in one launch: thread i writes its own seed
// entry, and a particle already holding the voxel has its entry rebuilt from
// its position and radius with the same operations (the same bits), so no
// thread reads an entry another is writing.
__device__ __forceinline__ float4 seed_entry(const float *__restrict__ pos,
                                             const float *__restrict__ rad, int i, float beta) {
  const float b = beta * rad[i];
  return make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], b * b);
}

template <bool UNIT_L>
__global__ void k_seed_fused(const float *__restrict__ pos, const float *__restrict__ rad,
                             float beta, float4 *__restrict__ seeds, int n, int res, float L,
                             float half_l, float voxel, int32_t *__restrict__ grid,
                             int *__restrict__ counters, int ncounters) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  // the pass counters of the fos_jfa_passes call that follows
  for (int t = i; t < ncounters; t += (int)(gridDim.x * blockDim.x)) counters[t] = 0;
  if (i >= n) return;
  const float4 si = seed_entry(pos, rad, i, beta);
  seeds[i] = si;
  int c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float nrm = (pos[3 * i + d] + half_l) / L;  // jfa.py:268-270
    c[d] = wrap_fast((int)floorf(nrm * (float)res), res);
  }
  const int64_t v = ((int64_t)c[0] * res + c[1]) * res + c[2];
  const float cx = vcenter(c[0], voxel, half_l), cy = vcenter(c[1], voxel, half_l),
              cz = vcenter(c[2], voxel, half_l);
  const float pi = power_dist<UNIT_L>(cx, cy, cz, si, L);
  int cur = __hip_atomic_load(&grid[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (true) {
    if (cur != -1) {
      float pc = power_dist<UNIT_L>(cx, cy, cz, seed_entry(pos, rad, cur, beta), L);
      bool better = (pi < pc) || (pi == pc && i < cur);
      if (!better) return;
    }
    int old = atomicCAS(&grid[v], cur, i);
    if (old == cur) return;
    cur = old;
  }
}

// @jfa-pass-kernel-begin (this region, the pass launch region and fos_common.h
// make the pass kernel's signature that counter profiles are matched by)
// ---------------------------------------------------------------------------
// One JFA pass (jfa.py:524-623)
// ---------------------------------------------------------------------------

// Whether pass p runs under the device-side early exit (jfa.py:1104-1112):
// it runs unless an earlier pass q >= min_passes-1 changed fewer than
// th * V voxels.  A pass that did not run counted no change (its counter is
// cleared before the schedule), and no change is below any positive th, so
// once one pass stops the schedule every later pass sees its predecessor
// stopped: the previous pass's count alone decides (one load, not p).
// Every thread derives the same answer.  Must be reached by whole waves (the
// striped totals use cross-lane sums).
__device__ __forceinline__ bool pass_active(const int *stripes, int p, int early,
                                            double th, int min_passes, double V) {
  if (!early || p < 1 || p - 1 < min_passes - 1) return true;
  return !((double)stripes_total(stripes, p - 1) / V < th);
}

// pass_active evaluated by wave 0 and shared with the block (all threads of
// the block must call it).
__device__ __forceinline__ bool block_pass_active(const int *stripes, int p, int early,
                                                  double th, int min_passes, double V) {
  if (!early) return true;
  __shared__ int s_active;
  if (threadIdx.x < 64) {
    const bool a = pass_active(stripes, p, early, th, min_passes, V);
    if (threadIdx.x == 0) s_active = a ? 1 : 0;
  }
  __syncthreads();
  return s_active != 0;
}

// Adds a workgroup's changed-voxel count (the popcounts of its waves'
// change ballots) to pass `pass`: an LDS sum and one atomic per workgroup, on
// the stripe of its block id.  A wave that ends on its own atomic keeps its
// slot until the atomic returns, which costs more than the two barriers.
// Every thread of the block must call it.
__device__ __forceinline__ void block_count(unsigned long long bal, int *stripes, int pass) {
  __shared__ int s_cnt;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  if ((threadIdx.x & 63) == 0 && bal) atomicAdd(&s_cnt, __popcll(bal));
  __syncthreads();
  if (threadIdx.x == 0 && s_cnt)
    atomicAdd(stripes + (pass * FOS_STRIPES + blockIdx.x % FOS_STRIPES) * FOS_STRIPE_INTS, s_cnt);
}

// One pass with per-voxel candidate lists.  The pass is bound by vector ALU
// issue: a wave pays for a candidate slot whenever any of its 64 voxels
// evaluates it, and with 27 fixed slots nearly every slot has some voxel
// that needs it.  So each voxel writes the owners it must evaluate to its
// own column of an LDS list (the incumbent first, then the skip-filtered
// candidates in scan order), and the wave walks the lists in steps of
// JFA_BATCH entries, so it runs max-over-lanes list length iterations instead
// of 27.  Lists are padded to a multiple of JFA_BATCH with their last entry:
// a repeated owner has the same distance and cannot win under '<'.
//
// Skip rule: an owner equal to the incumbent or to the mirror candidate on
// the -1 side of an axis where this one sits at +1 (labels after a pass of
// step s repeat with period 2s along each axis); NEIGH adds the -1 neighbour
// along each axis, which pays once the field has converged (the repeated
// step-1 pass).  Both compare only with earlier candidates.
#define JFA_BATCH 2
// list rows: 27 owners plus the B - 1 entries a batch of B reads past the end
#define JFA_LIST_ROWS(B) (27 + (B) - 1)
#define JFA_LIST JFA_LIST_ROWS(JFA_BATCH)

template <bool NEIGH>
__device__ __forceinline__ bool jfa_skip(const int lab[27], int q, int inc) {
  const int dx = q / 9, dy = (q / 3) % 3, dz = q % 3;
  const int c = lab[q];
  bool skip = c < 0 || c == inc;
  if (dz == 2) skip |= c == lab[q - 2];
  if (dy == 2) skip |= c == lab[q - 6];
  if (dx == 2) skip |= c == lab[q - 18];
  if (NEIGH) {
    if (dz >= 1) skip |= c == lab[q - 1];
    if (dy >= 1) skip |= c == lab[q - 3];
    if (dx >= 1) skip |= c == lab[q - 9];
  }
  return skip;
}

// Pass flavours of the list kernel.
#define JFA_PLAIN 0    // mirror skip rule
#define JFA_SETTLED 1  // also skip repeats of the -1 neighbours
#define JFA_HALF 3     // 2*step == res: the -step and +step candidates coincide

// The owner of one voxel from its 27 candidate labels (lab[13] = current
// owner): the kept candidates go to this thread's LDS list column, then the
// list is walked BATCH entries at a time in list order with strict '<'.
// LT: the LDS list entry type -- int, or unsigned short when every owner id
// minus `base` is below 65536 (half the LDS, more resident workgroups).
// DIN: bd holds the incumbent's power distance on entry (the previous
// pass's result for this voxel: the same voxel centre and seed, so the same
// fp32 value), and the incumbent is not evaluated again.  On return bd is
// the winner's power distance (1e10 when no owner).
template <bool UNIT_L, int MODE, typename LT = int, int BATCH = JFA_BATCH, bool DIN = false>
__device__ __forceinline__ int jfa_list_eval(const int lab[27], float cx, float cy, float cz,
                                             const float4 *__restrict__ seeds, float L,
                                             LT (*s_list)[256], int base, float &bd) {
  // This thread's list column.  Each thread owns its column, so any
  // one-to-one column map is correct.  With 16-bit entries two columns share
  // a dword; lanes l and l + 32 of a wave take the two halves of one dword
  // (lane l's column 2l, lane l + 32's 2l + 1), so lanes that append at
  // different rows in the same half-wave never meet in one bank.
  const int tid = (int)threadIdx.x;
  const int t = sizeof(LT) == 2 ? (tid & ~63) | ((tid & 31) << 1) | ((tid >> 5) & 1) : tid;
  const int inc = lab[13];
  const bool have = DIN && inc >= 0;
  // Branch-free list build: every candidate is written at the current
  // end and the end advances only for kept owners.
  s_list[0][t] = (LT)(inc - base);
  int cnt = (inc >= 0 && !have) ? 1 : 0;
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    if (q == 13) continue;
    s_list[cnt][t] = (LT)(lab[q] - base);
    cnt += jfa_skip<MODE != JFA_PLAIN && MODE != JFA_HALF>(lab, q, inc) ? 0 : 1;
  }
  int best = have ? inc : -1;
  if (!have) bd = 1e10f;
  if (cnt > 0) {
    // entries past the end repeat the first kept owner (a valid id; a
    // repeat never wins under the strict '<')
    const LT first = s_list[0][t];
#pragma unroll
    for (int i = 0; i < BATCH - 1; ++i) s_list[cnt + i][t] = first;
    for (int k = 0; k < cnt; k += BATCH) {
      int c[BATCH];
      float4 sd[BATCH];
#pragma unroll
      for (int i = 0; i < BATCH; ++i) c[i] = (int)s_list[k + i][t] + base;
#pragma unroll
      for (int i = 0; i < BATCH; ++i)
        sd[i] = *(const float4 *)((const char *)seeds + ((uint32_t)c[i] << 4));
#pragma unroll
      for (int i = 0; i < BATCH; ++i) {
        const float pd = power_dist<UNIT_L>(cx, cy, cz, sd[i], L);
        if (pd < bd) {
          bd = pd;
          best = c[i];
        }
      }
    }
  }
  return best;
}

// Plane order of a pass of step `step`: the i-th x-plane processed is the
// k-th of residue class r (x = r + k*step), classes in turn.  Consecutive
// planes of a class share two of their three source planes (x - step, x,
// x + step), so each source plane is read about once while it is in the
// L2 of the XCD that processes the class, instead of again `step` planes
// later.  (Class r < res % step holds res / step + 1 planes, the rest
// res / step.)  The divisions by q = res / step >= 2 and q + 1 use the
// host's magic numbers mq = ceil(2^32 / q), mq1 = ceil(2^32 / (q + 1)):
// (i * m) >> 32 is the exact quotient for every i < 8192.
__device__ __forceinline__ int jfa_plane(int i, int step, int qr, unsigned mq, unsigned mq1) {
  const int q = qr >> 16, rem = qr & 0xffff;
  const int big = rem * (q + 1);
  if (i < big) {
    const int r = (int)__umulhi((unsigned)i, mq1);
    return r + (i - r * (q + 1)) * step;
  }
  const int j = i - big;
  const int r = (int)__umulhi((unsigned)j, mq);
  return rem + r + (j - r * q) * step;
}

// One tile (4 z-rows of 64 voxels at one x, this thread's wave one row) of a
// list pass; returns this wave's change ballot (0 for an idle row).  The
// caller has checked that the pass is active.  CLASS: tiles take their
// x-planes in jfa_plane order.
// DM: 0 no distance buffers; 1 writes each owned voxel's winning power
// distance to dist_out; 2 also starts from dist_in (the previous pass's
// dist_out, read only where that pass left an owner).
template <bool UNIT_L, int MODE, typename LT = int, int BATCH = JFA_BATCH, int DM = 0,
          bool CLASS = false>
__device__ __forceinline__ unsigned long long jfa_list_tile(
    const float4 *__restrict__ seeds, int res, float L, float half_l, float voxel, int step,
    const int32_t *__restrict__ in, int32_t *__restrict__ out, int tile, int nyb, int nzb,
    int ntiles, LT (*s_list)[256], int base = 0,
    const float *__restrict__ dist_in = nullptr, float *__restrict__ dist_out = nullptr,
    int qr = 0, unsigned mq = 0, unsigned mq1 = 0) {
  const int t = (int)threadIdx.x, lane = t & 63;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  int changed = 0;
  if (tile < ntiles) {
    const int zb = tile % nzb;
    const int t2 = tile / nzb;
    const int yb = t2 % nyb;
    const int x = CLASS ? jfa_plane(t2 / nyb, step, qr, mq, mq1) : t2 / nyb;
    const int y = yb * 4 + w;
    const int z = zb * 64 + lane;
    const bool work = y < res && z < res;
    const size_t v = ((size_t)x * res + y) * res + z;
    if (work) {
      const float cx = vcenter(x, voxel, half_l), cy = vcenter(y, voxel, half_l),
                  cz = vcenter(z, voxel, half_l);
      float bd = DM == 2 ? dist_in[v] : 0.0f;
      // x, y are wave-uniform: the nine row bases live in scalar registers
      // and each label load adds one of three per-lane 32-bit z offsets
      // (scalar base + vector offset addressing; res <= 1290 keeps a row
      // under 2^32 B).
      uint32_t zoff[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) zoff[c] = (uint32_t)wrap1(z + (c - 1) * step, res) << 2;
      int lab[27];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          const int xr = wrap1(x + (a - 1) * step, res), yr = wrap1(y + (b - 1) * step, res);
          const char *row = (const char *)(in + ((size_t)xr * res + yr) * res);
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const int q = (a * 3 + b) * 3 + c;
            if (MODE == JFA_HALF && a == 2)
              lab[q] = lab[q - 18];
            else if (MODE == JFA_HALF && b == 2)
              lab[q] = lab[q - 6];
            else if (MODE == JFA_HALF && c == 2)
              lab[q] = lab[q - 2];
            else
              lab[q] = *(const int32_t *)(row + zoff[c]);
          }
        }
      const int best = jfa_list_eval<UNIT_L, MODE, LT, BATCH, DM == 2>(lab, cx, cy, cz, seeds, L,
                                                                         s_list, base, bd);
      out[v] = best;
      // a voxel with no owner has no distance to carry: the next pass reads
      // dist_in only under an owned incumbent
      if (DM && best >= 0) dist_out[v] = bd;
      changed = (best != lab[13]);
    }
  }
  return __ballot(changed);
}

// One flat tile of a list pass: 256 consecutive voxels in (y, z) order of
// one x-plane (tile = x * tpp + block of the plane, tpp = ceil(res^2/256)),
// one per thread.  A wave may span two rows, so the row offsets are per
// lane; no lanes idle at the end of a row whatever res is (only at the end
// of the plane).  Returns this wave's change ballot.
template <bool UNIT_L, int MODE, typename LT = int>
__device__ __forceinline__ unsigned long long jfa_flat_tile(
    const float4 *__restrict__ seeds, int res, float L, float half_l, float voxel, int step,
    const int32_t *__restrict__ in, int32_t *__restrict__ out, int tile, int tpp,
    LT (*s_list)[256], int base = 0) {
  const int t = (int)threadIdx.x;
  const int x = tile / tpp;
  const int r = (tile - x * tpp) * 256 + t;
  int changed = 0;
  if (x < res && r < res * res) {
    const int y = r / res, z = r - y * res;
    const size_t v = ((size_t)x * res + y) * res + z;
    const float cx = vcenter(x, voxel, half_l), cy = vcenter(y, voxel, half_l),
                cz = vcenter(z, voxel, half_l);
    uint32_t yo[3], zo[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      yo[c] = (uint32_t)wrap1(y + (c - 1) * step, res) * (uint32_t)res;
      zo[c] = (uint32_t)wrap1(z + (c - 1) * step, res);
    }
    int lab[27];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int32_t *plane = in + (size_t)wrap1(x + (a - 1) * step, res) * res * res;
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const int q = (a * 3 + b) * 3 + c;
          if (MODE == JFA_HALF && a == 2)
            lab[q] = lab[q - 18];
          else if (MODE == JFA_HALF && b == 2)
            lab[q] = lab[q - 6];
          else if (MODE == JFA_HALF && c == 2)
            lab[q] = lab[q - 2];
          else
            lab[q] = plane[yo[b] + zo[c]];
        }
    }
    float bd;
    const int best = jfa_list_eval<UNIT_L, MODE, LT>(lab, cx, cy, cz, seeds, L, s_list, base, bd);
    out[v] = best;
    changed = (best != lab[13]);
  }
  return __ballot(changed);
}

template <bool UNIT_L, int MODE, typename LT, int BATCH = JFA_BATCH, int DM = 0,
          bool CLASS = false>
__global__ void __launch_bounds__(256)
    k_jfa_pass_list(const float4 *__restrict__ seeds, int res, float L, float half_l,
                    float voxel, int step, const int32_t *__restrict__ in,
                    int32_t *__restrict__ out, int *__restrict__ stripes, int pass, int early,
                    double th, int min_passes, int nyb, int nzb, int ntiles, int per_xcd,
                    const float *__restrict__ dist_in, float *__restrict__ dist_out, int qr,
                    unsigned mq, unsigned mq1) {
  __shared__ LT s_list[JFA_LIST_ROWS(BATCH)][256];
  if (!block_pass_active(stripes, pass, early, th, min_passes, (double)res * res * res)) return;
  const int tile = xcd_tile(blockIdx.x, per_xcd);
  const unsigned long long bal = jfa_list_tile<UNIT_L, MODE, LT, BATCH, DM, CLASS>(
      seeds, res, L, half_l, voxel, step, in, out, tile, nyb, nzb, ntiles, s_list, 0, dist_in,
      dist_out, qr, mq, mq1);
  block_count(bal, stripes, pass);
}

// The pass at step = res / 2 (jfa.py:524-623 with 2 * step == res).  Then
// -step and +step are the same voxel, so voxel v's 27 candidates sit on the
// 8 voxels of its orbit {v + step * e : e in {0,1}^3}, and every voxel of an
// orbit has the same 8 candidates.  One thread takes one orbit: it loads the
// orbit's 8 labels once and decides its 8 voxels.  For the voxel at orbit
// element e, candidate (a, b, c) of the scan sits at element e ^ m with m =
// (a != 1, b != 1, c != 1), so the candidates' first occurrences in scan
// order are e ^ 7, e ^ 6, ..., e ^ 1, e ^ 0 = the voxel itself: the
// incumbent, then those seven in that order, strict '<' (a repeat never
// wins).  Right after the seeding most orbits hold no owner at all, so the
// pass costs about its memory traffic.  DM: 1 also writes each owned
// voxel's winning power distance.
template <bool UNIT_L, int DM>
__global__ void __launch_bounds__(256)
    k_jfa_half(const float4 *__restrict__ seeds, int res, float L, float half_l, float voxel,
               const int32_t *__restrict__ in, int32_t *__restrict__ out,
               int *__restrict__ stripes, int pass, int early, double th, int min_passes,
               float *__restrict__ dist_out) {
  if (!block_pass_active(stripes, pass, early, th, min_passes, (double)res * res * res)) return;
  const int s = res >> 1;
  const int64_t orbits = (int64_t)s * s * s;
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int changed = 0;
  if (o < orbits) {
    const int oz = (int)(o % s), oy = (int)((o / s) % s), ox = (int)(o / ((int64_t)s * s));
    int lab[8];
    size_t vi[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int x = ox + ((e >> 2) & 1) * s, y = oy + ((e >> 1) & 1) * s, z = oz + (e & 1) * s;
      vi[e] = ((size_t)x * res + y) * res + z;
      lab[e] = in[vi[e]];
    }
    int any = 0;
#pragma unroll
    for (int e = 0; e < 8; ++e) any |= lab[e] >= 0;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      int best = lab[e];
      float bd = 1e10f;
      if (any) {
        const int x = ox + ((e >> 2) & 1) * s, y = oy + ((e >> 1) & 1) * s, z = oz + (e & 1) * s;
        const float cx = vcenter(x, voxel, half_l), cy = vcenter(y, voxel, half_l),
                    cz = vcenter(z, voxel, half_l);
        if (best >= 0) bd = power_dist<UNIT_L>(cx, cy, cz, seeds[best], L);
#pragma unroll
        for (int m = 7; m >= 1; --m) {
          const int c = lab[e ^ m];
          if (c >= 0 && c != best) {
            const float pd = power_dist<UNIT_L>(cx, cy, cz, seeds[c], L);
            if (pd < bd) {
              bd = pd;
              best = c;
            }
          }
        }
      }
      out[vi[e]] = best;
      if (DM && best >= 0) dist_out[vi[e]] = bd;
      changed += best != lab[e];
    }
  }
  // the block's changed voxels, one atomic per block (block_count's sum)
  __shared__ int s_cnt;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  int w = changed;
  for (int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off, 64);
  if ((threadIdx.x & 63) == 0 && w) atomicAdd(&s_cnt, w);
  __syncthreads();
  if (threadIdx.x == 0 && s_cnt)
    atomicAdd(stripes + (pass * FOS_STRIPES + blockIdx.x % FOS_STRIPES) * FOS_STRIPE_INTS, s_cnt);
}

// @jfa-pass-kernel-end

// Ends the pass loop: records passes executed and the exit flag, and copies
// temp into grid when an odd number of passes ran (jfa.py:1115-1116).
// The exit flag is set when the condition held after the last executed pass,
// including the final scheduled one (jfa.py:1109-1112 breaks there too).
__global__ void k_jfa_finish(int res, int num_passes, int early, double th,
                             int min_passes, const int *__restrict__ stripes,
                             int32_t *__restrict__ info, const int32_t *__restrict__ temp,
                             int32_t *__restrict__ grid) {
  const double V = (double)res * res * res;
  __shared__ int s_tot[FOS_PASS_SLOTS];
  __shared__ int s_exec;
  // thread p sums pass p's stripes (independent loads, not one dependent
  // chain per pass); thread 0 then applies pass_active's rule to the totals
  for (int p = (int)threadIdx.x; p < num_passes; p += (int)blockDim.x) {
    int c = 0;
    for (int q = 0; q < FOS_STRIPES; ++q) c += stripes[(p * FOS_STRIPES + q) * FOS_STRIPE_INTS];
    s_tot[p] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    auto active = [&](int p) {
      return !early || p < 1 || p - 1 < min_passes - 1 || !((double)s_tot[p - 1] / V < th);
    };
    int executed = 0;
    while (executed < num_passes && active(executed)) ++executed;
    if (blockIdx.x == 0) {
      for (int p = 0; p < num_passes; ++p) info[2 + p] = p < executed ? s_tot[p] : 0;
      info[0] = executed;
      info[1] = (early && executed > 0 && !active(executed)) ? 1 : 0;
    }
    s_exec = executed;
  }
  __syncthreads();
  const int executed = s_exec;
  const int64_t total = (int64_t)res * res * res;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // The last executed pass also decides the exit when it is the final one.
  if (executed & 1) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = t; i < total; i += stride) grid[i] = temp[i];
  }
}

// Adds the striped total of one pass slot into *dst (one wave).
__global__ void k_stripes_add(const int *__restrict__ stripes, int slot, int32_t *dst) {
  const int t = stripes_total(stripes, slot);
  if (threadIdx.x == 0) *dst += t;
}

__global__ void k_copy_i32(const int32_t *__restrict__ s, int32_t *__restrict__ d,
                           int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) d[i] = s[i];
}

// ---------------------------------------------------------------------------
// Majority filter (jfa.py:640-699)
// ---------------------------------------------------------------------------

// This thread's voxel under the XCD-aware tile order (tile = 1 x-plane x 4
// y-rows x 64 z, x-major); false for padding threads.
__device__ __forceinline__ bool vox_of(int res, int nyb, int nzb, int ntiles, int per,
                                       int &x, int &y, int &z) {
  const int tile = xcd_tile(blockIdx.x, per);
  if (tile >= ntiles) return false;
  const int zb = tile % nzb, t2 = tile / nzb;
  const int yb = t2 % nyb;
  x = t2 / nyb;
  z = zb * 64 + (int)(threadIdx.x & 63);
  y = yb * 4 + (int)(threadIdx.x >> 6);
  return z < res && y < res;
}

// The 6 face neighbours in reference order: axis 0..2, direction -1 then +1.
__device__ __forceinline__ void face_neighbours(const int32_t *__restrict__ g, int res, int x,
                                                int y, int z, int nb[6]) {
  const int rr = res * res;
  nb[0] = g[wrap1(x - 1, res) * rr + y * res + z];
  nb[1] = g[wrap1(x + 1, res) * rr + y * res + z];
  nb[2] = g[x * rr + wrap1(y - 1, res) * res + z];
  nb[3] = g[x * rr + wrap1(y + 1, res) * res + z];
  nb[4] = g[x * rr + y * res + wrap1(z - 1, res)];
  nb[5] = g[x * rr + y * res + wrap1(z + 1, res)];
}

__device__ __forceinline__ void majority_vox(int res, const int32_t *__restrict__ g,
                                             int32_t *__restrict__ f, int x, int y, int z) {
  const int v = (x * res + y) * res + z;
  const int my = g[v];
  if (my < 0) {
    f[v] = -1;
    return;
  }
  int t[7];
  t[0] = my;
  face_neighbours(g, res, x, y, z, t + 1);
  // Tally in first-seen order (centre, then the neighbours in order, -1
  // skipped); a label's count is its number of occurrences among the seven
  // (the centre counts once).  The first label with the strictly largest
  // count wins.  Each of the 21 pairs is compared once: an equal pair adds
  // to both counts and marks the later entry as a repeat.  Static indices
  // only, so the tally stays in registers.
  int c[7];
  bool rep[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    c[k] = 1;
    rep[k] = t[k] < 0;
  }
#pragma unroll
  for (int k = 1; k < 7; ++k)
#pragma unroll
    for (int j = 0; j < k; ++j) {
      const bool e = t[j] == t[k];
      c[j] += e ? 1 : 0;
      c[k] += e ? 1 : 0;
      rep[k] |= e;
    }
  int maxc = 0, best = my;
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    if (!rep[k] && c[k] > maxc) {
      maxc = c[k];
      best = t[k];
    }
  }
  f[v] = best;
}

__global__ void __launch_bounds__(256)
    k_majority(int res, const int32_t *__restrict__ g, int32_t *__restrict__ f, int nyb,
               int nzb, int ntiles, int per) {
  int x, y, z;
  if (!vox_of(res, nyb, nzb, ntiles, per, x, y, z)) return;
  majority_vox(res, g, f, x, y, z);
}

// ---------------------------------------------------------------------------
// Face-pair accumulation (jfa.py:703-762), serial semantics:
//   row a lists its partners b > a by first-encounter key
//   k = 6*lin(I) + 2*axis + (dir > 0), first 32 kept, counts complete.
// Emission inserts (b, min key, count) into a per-row open-addressed table;
// a compaction pass ranks each row's entries by key into face_ids/counts.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void tab_insert(int *__restrict__ tb,
                                           unsigned *__restrict__ tk,
                                           int *__restrict__ tc, int cap,
                                           int a, int b, unsigned key, int cnt,
                                           int *overflow) {
  const int64_t base = (int64_t)a * cap;
  const unsigned h = (unsigned)b * 2654435761u;
  for (int p = 0; p < cap; ++p) {
    const int s = (int)((h + (unsigned)p) & (unsigned)(cap - 1));
    int cur = __hip_atomic_load(&tb[base + s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == -1) {
      int old = atomicCAS(&tb[base + s], -1, b);
      cur = (old == -1) ? b : old;
    }
    if (cur == b) {
      atomicMin(&tk[base + s], key);
      atomicAdd(&tc[base + s], cnt);
      return;
    }
  }
  atomicOr(overflow, 1);
}

__device__ __forceinline__ int wave_sum(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ unsigned wave_min_u(unsigned v) {
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned w = (unsigned)__shfl_xor((int)v, o, 64);
    v = w < v ? w : v;
  }
  return v;
}

// A voxel first merges its repeated partners (one entry per distinct b with
// the smallest key and the count).  The wave then merges equal (a, b) pairs
// across lanes -- neighbouring voxels on one boundary emit the same pair --
// so each distinct pair costs one table insert per wave.  Counts add and keys
// take the minimum, so the merge order does not matter.
__global__ void __launch_bounds__(256)
    k_face_emit(int res, int n, const int32_t *__restrict__ f, int *tb, unsigned *tk,
                int *tc, int cap, int *overflow, int nyb, int nzb, int ntiles, int per) {
  int x = 0, y = 0, z = 0;
  const bool live = vox_of(res, nyb, nzb, ntiles, per, x, y, z);
  const int lane = (int)(threadIdx.x & 63);
  int a = -1;
  int nb[6] = {-1, -1, -1, -1, -1, -1};
  unsigned kbase = 0;
  if (live) {
    const int v = (x * res + y) * res + z;
    a = f[v];
    if (a >= 0 && a < n) {
      face_neighbours(f, res, x, y, z, nb);
      kbase = 6u * (unsigned)v;
    } else {
      a = -1;
    }
  }
  bool own[6];
  int oc[6];
  unsigned okey[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int b = nb[k];
    bool has = a >= 0 && b >= 0 && b < n && a < b;
    int cnt = 1;
#pragma unroll
    for (int q = 0; q < k; ++q)
      if (nb[q] == b) has = false;
#pragma unroll
    for (int q = k + 1; q < 6; ++q)
      if (nb[q] == b) cnt += 1;
    const unsigned key = kbase + (unsigned)k;
    // merge equal pairs; each distinct pair's leader lane keeps the totals
    own[k] = false;
    oc[k] = 0;
    okey[k] = 0;
    unsigned long long pend = __ballot(has);
    while (pend) {
      const int leader = __ffsll((long long)pend) - 1;
      const int la = __shfl(a, leader, 64), lb = __shfl(b, leader, 64);
      const bool m = has && a == la && b == lb;
      const int c = wave_sum(m ? cnt : 0);
      const unsigned kk = wave_min_u(m ? key : 0xffffffffu);
      if (lane == leader) {
        own[k] = true;
        oc[k] = c;
        okey[k] = kk;
      }
      if (m) has = false;
      pend = __ballot(has);
    }
  }
  // Inserts: first probes of all owned pairs issued together, then resolved.
  const int64_t base = (int64_t)(a < 0 ? 0 : a) * cap;
  int probe[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    probe[k] = -2;
    if (own[k]) {
      const int s = (int)(((unsigned)nb[k] * 2654435761u) & (unsigned)(cap - 1));
      probe[k] = tb[base + s];
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    if (!own[k]) continue;
    const int b = nb[k];
    if (probe[k] == b) {
      const int s = (int)(((unsigned)b * 2654435761u) & (unsigned)(cap - 1));
      atomicMin(&tk[base + s], okey[k]);
      atomicAdd(&tc[base + s], oc[k]);
    } else {
      tab_insert(tb, tk, tc, cap, a, b, okey[k], oc[k], overflow);
    }
  }
}

// Workgroup-level merge: a tile of TX x-planes x (4 * 64 / ZW) y-rows x ZW
// z-voxels (each wave covers 64 / ZW rows of ZW voxels) merges its pairs in
// an LDS hash keyed by (a, b) (count sum, key minimum) and then inserts each
// distinct pair once into the row tables.  Squarer tiles (smaller ZW) cross
// fewer cell boundaries per voxel, so fewer pairs reach the row tables.  A pair that finds no
// LDS slot within EMIT_PROBES probes goes straight to the row table, so the
// LDS table can be smaller than the worst case.  512 slots (8 KB of LDS with
// the keys and counts): a tile of 8 x 8 x 32 voxels holds a few dozen distinct
// pairs, and the table's clear and flush loops and its LDS footprint shrink
// with it (2048 slots measured 12% slower at C2, 55% at C5; 256 overflowed
// into the row tables on the deeper 16-plane tiles).
#define EMIT_HS 512
#define EMIT_PROBES 32
#define EMIT_EMPTY 0xffffffffffffffffull

__device__ __forceinline__ unsigned pair_hash(int a, int b) {
  return ((unsigned)a * 0x9E3779B1u) ^ ((unsigned)b * 0x85EBCA77u);
}

// One emission tile (all threads of the workgroup call it); nyb, nzb count
// the tile's y and z blocks (y block 4 * 64 / ZW rows, z block ZW voxels).
template <int TX, int ZW = 64>
__device__ __forceinline__ void face_emit_tile(int tile, int res, int n,
                                               const int32_t *__restrict__ f, int *tb,
                                               unsigned *tk, int *tc, int cap, int *overflow,
                                               int nyb, int nzb, int ntiles,
                                               unsigned long long *s_pair, unsigned *s_key,
                                               int *s_cnt) {
  const int t = (int)threadIdx.x;
  for (int h = t; h < EMIT_HS; h += 256) {
    s_pair[h] = EMIT_EMPTY;
    s_key[h] = 0xffffffffu;
    s_cnt[h] = 0;
  }
  __syncthreads();
  if (tile < ntiles) {
    const int zb = tile % nzb, t2 = tile / nzb;
    const int yb = t2 % nyb, x0 = (t2 / nyb) * TX;
    const int z = zb * ZW + (t % ZW), y = yb * (256 / ZW) + t / ZW;
    const bool in_yz = y < res && z < res;
    // the next plane's seven labels are loaded before this plane's pairs go
    // into the hash, so their latency overlaps the LDS work
    int a_nx = -1, nb_nx[6];
    if (in_yz && x0 < res) {
      a_nx = f[(x0 * res + y) * res + z];
      face_neighbours(f, res, x0, y, z, nb_nx);
    }
    for (int xt = 0; xt < TX; ++xt) {
      const int x = x0 + xt;
      if (x >= res || !in_yz) break;
      const int v = (x * res + y) * res + z;
      const int a = a_nx;
      int nb[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) nb[k] = nb_nx[k];
      if (xt + 1 < TX && x + 1 < res) {
        a_nx = f[v + res * res];
        face_neighbours(f, res, x + 1, y, z, nb_nx);
      }
      if (a < 0 || a >= n) continue;
      const unsigned kbase = 6u * (unsigned)v;
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const int b = nb[k];
        bool has = b >= 0 && b < n && a < b;
        int cnt = 1;
#pragma unroll
        for (int q = 0; q < k; ++q)
          if (nb[q] == b) has = false;
#pragma unroll
        for (int q = k + 1; q < 6; ++q)
          if (nb[q] == b) cnt += 1;
        if (!has) continue;
        const unsigned key = kbase + (unsigned)k;
        const unsigned long long pk = ((unsigned long long)(unsigned)a << 32) | (unsigned)b;
        unsigned h = pair_hash(a, b) & (EMIT_HS - 1);
        int slot = -1;
        for (int p = 0; p < EMIT_PROBES; ++p) {
          unsigned long long cur = s_pair[h];
          if (cur == EMIT_EMPTY) {
            cur = atomicCAS(&s_pair[h], EMIT_EMPTY, pk);
            if (cur == EMIT_EMPTY) cur = pk;
          }
          if (cur == pk) {
            slot = (int)h;
            break;
          }
          h = (h + 1) & (EMIT_HS - 1);
        }
        if (slot >= 0) {
          atomicMin(&s_key[slot], key);
          atomicAdd(&s_cnt[slot], cnt);
        } else {
          tab_insert(tb, tk, tc, cap, a, b, key, cnt, overflow);
        }
      }
    }
  }
  __syncthreads();
  for (int h = t; h < EMIT_HS; h += 256) {
    const unsigned long long pk = s_pair[h];
    if (pk == EMIT_EMPTY) continue;
    const int a = (int)(pk >> 32), b = (int)(pk & 0xffffffffu);
    const int64_t base = (int64_t)a * cap;
    const int s = (int)(((unsigned)b * 2654435761u) & (unsigned)(cap - 1));
    if (tb[base + s] == b) {
      atomicMin(&tk[base + s], s_key[h]);
      atomicAdd(&tc[base + s], s_cnt[h]);
    } else {
      tab_insert(tb, tk, tc, cap, a, b, s_key[h], s_cnt[h], overflow);
    }
  }
}

template <int TX, int ZW>
__global__ void __launch_bounds__(256)
    k_face_emit_lds(int res, int n, const int32_t *__restrict__ f, int *tb, unsigned *tk,
                    int *tc, int cap, int *overflow, int nyb, int nzb, int ntiles, int per) {
  __shared__ unsigned long long s_pair[EMIT_HS];
  __shared__ unsigned s_key[EMIT_HS];
  __shared__ int s_cnt[EMIT_HS];
  face_emit_tile<TX, ZW>(xcd_tile(blockIdx.x, per), res, n, f, tb, tk, tc, cap, overflow, nyb, nzb,
                     ntiles, s_pair, s_key, s_cnt);
}

// One wave per row: rank entries by key (keys are unique within a row), write
// the first 32 into face_ids/face_counts, clear the rest, reset the table.
// in_count (may be NULL): also clears what finalize accumulates into (as
// k_fin_clear does), for a finalize that follows at once.
__global__ void __launch_bounds__(256)
    k_face_compact(int n, int *tb, unsigned *tk, int *tc, int cap,
                   int32_t *__restrict__ face_ids,
                   int32_t *__restrict__ face_counts, int *__restrict__ in_count,
                   int *__restrict__ in_fill, unsigned *__restrict__ dec,
                   int *__restrict__ chg, int nchg, int *__restrict__ dirty0,
                   int *__restrict__ dirty1, int *__restrict__ zero_cnt) {
  extern __shared__ unsigned s_keys[];
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + wid;
  if (zero_cnt && blockIdx.x == 0 && threadIdx.x == 0) *zero_cnt = 0;
  if (in_count) {
    if (row < n && lane == 0) {
      in_count[row] = 0;
      in_fill[row] = 0;
      dec[row] = 0u;
      dirty0[row] = 0;
      dirty1[row] = 0;
    }
    if (blockIdx.x == 0) {
      if (threadIdx.x == 0) in_count[n] = 0;
      for (int t = (int)threadIdx.x; t < nchg; t += 256) chg[t] = 0;
    }
  }
  unsigned *keys = s_keys + wid * cap;
  const bool live = row < n;
  const int64_t base = (int64_t)row * cap;
  int nvalid = 0;
  for (int e = lane; e < cap; e += 64) {
    unsigned k = 0xffffffffu;
    if (live && tb[base + e] != -1) {
      k = tk[base + e];
      nvalid += 1;
    }
    keys[e] = k;
  }
  for (int o = 32; o > 0; o >>= 1) nvalid += __shfl_xor(nvalid, o, 64);
  __syncthreads();
  if (!live) return;
  for (int e = lane; e < cap; e += 64) {
    const int b = tb[base + e];
    if (b == -1) continue;
    const unsigned k = keys[e];
    int rank = 0;
    for (int q = 0; q < cap; ++q) rank += (keys[q] < k) ? 1 : 0;
    if (rank < FOS_MAX_NEIGHBORS) {
      face_ids[(int64_t)row * FOS_MAX_NEIGHBORS + rank] = b;
      face_counts[(int64_t)row * FOS_MAX_NEIGHBORS + rank] = tc[base + e];
    }
    tb[base + e] = -1;
    tk[base + e] = 0xffffffffu;
    tc[base + e] = 0;
  }
  if (lane < FOS_MAX_NEIGHBORS && lane >= nvalid) {
    face_ids[(int64_t)row * FOS_MAX_NEIGHBORS + lane] = -1;
    face_counts[(int64_t)row * FOS_MAX_NEIGHBORS + lane] = 0;
  }
}

// ---------------------------------------------------------------------------
// Finalize (jfa.py:789-857, add_neighbor jfa.py:400-443)
//
// Serial order: rows i ascending, slots s ascending; an accepted pair (i, j)
// appends j to row i and i to row j (while the row holds < 32).  Row i's
// hysteresis read prev[i][s] >= T_DROP is true when position s was written
// this frame before the read -- i.e. s < used_i at that moment, where
// used_i = min(32, L_i + own accepts so far) and L_i = accepted appends from
// rows i' < i -- or when the previous frame left prev[i][s] >= T_DROP.
// Accept writes are scores >= T_DROP, reject writes land at the slot's own
// index (never read again this frame).  Rows depend only on lower rows, so
// iterating "decide rows given L, recount L" from L = 0 reaches the serial
// result (monotone, unique fixed point).
// ---------------------------------------------------------------------------

#define CLS_INVALID 0
#define CLS_REJECT 1
#define CLS_ACCEPT 2
#define CLS_AMBIG 3
#define CLS_OLD 4

// jfa.py:379-397 expected_area
__device__ __forceinline__ float expected_area(float ri, float rj, float voxel) {
  float r_mean = 0.5f * (ri + rj);
  float t = (2.0f * r_mean) / voxel;
  return 1.5f * (t * t);
}

// pvox (may be NULL): per-particle voxel edge (a batch of foams), else voxel.
__global__ void k_fin_classify(int n, const float *__restrict__ rad, float voxel,
                               const float *__restrict__ pvox,
                               const int32_t *__restrict__ face_ids,
                               const int32_t *__restrict__ face_counts,
                               const float *__restrict__ prev,
                               float *__restrict__ score,
                               unsigned char *__restrict__ cls,
                               int *__restrict__ in_count) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * FOS_MAX_NEIGHBORS) return;
  const int i = (int)(t / FOS_MAX_NEIGHBORS);
  const int j = face_ids[t];
  unsigned char c = CLS_INVALID;
  float sc = 0.0f;
  if (j >= 0 && j < n) {
    const int v = face_counts[t];
    if (v > 0) {
      const float e = expected_area(rad[i], rad[j], pvox ? pvox[i] : voxel);
      if (e > 0.0f) {
        sc = (float)v / e;
        if (sc >= 0.15f)
          c = CLS_ACCEPT;
        else if (sc < 0.10f)
          c = CLS_REJECT;
        else
          c = CLS_AMBIG;
        if (prev[t] >= 0.10f) c |= CLS_OLD;
        atomicAdd(&in_count[j], 1);
      }
    }
  }
  score[t] = sc;
  cls[t] = c;
}

__global__ void k_fin_scatter(int n, const int32_t *__restrict__ face_ids,
                              const unsigned char *__restrict__ cls,
                              const int *__restrict__ in_start,
                              int *__restrict__ in_fill, int *__restrict__ in_src) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * FOS_MAX_NEIGHBORS) return;
  if ((cls[t] & 3) == CLS_INVALID) return;
  const int j = face_ids[t];
  const int p = in_start[j] + atomicAdd(&in_fill[j], 1);
  in_src[p] = (int)t;
}

// Places each incoming record at its rank within its row's list (ascending
// source row, then slot) so the materialised neighbour rows list lower
// partners in row order.  Lists are short; each record counts the smaller
// ones in its segment.
__global__ void k_fin_rank(int n, const int32_t *__restrict__ face_ids,
                           const int *__restrict__ in_start, const int *__restrict__ in_src,
                           int *__restrict__ in_sorted) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= in_start[n]) return;
  const int src = in_src[p];
  const int j = face_ids[src];
  const int b = in_start[j], e = in_start[j + 1];
  int rank = 0;
  for (int q = b; q < e; ++q) rank += in_src[q] < src ? 1 : 0;
  in_sorted[b + rank] = src;
}

// One fixed-point round for row i: L_i = min(32, accepted appends from lower
// rows under the previous round's decisions), then the row's decisions in
// slot order (an ambiguous score is kept when its position was written this
// frame -- s < used -- or last frame left it >= T_DROP).  Decisions are
// packed as one bit per slot.
// 8 lanes per row gather the accepted-incoming count (an order-free sum);
// lane 0 then decides the row.
// Each round keeps one change flag: the round sets chg_cur when a row's
// decisions change.  chg_prev (may be NULL), the previous round's flag,
// still clear means that round changed nothing -- the fixed point -- and the
// round does nothing (so its own flag stays clear and the skip cascades).
// failed (may be NULL; given to the last round of a fixed count) is set when
// that round changed anything.
// Dirty rows: a row's decisions depend only on the bits of its incoming
// slots, so a row none of whose incoming bits changed in the previous round
// keeps its decisions.  The previous round marked the targets of the bits it
// changed in dirty_in (face_ids[row][slot] of each changed slot); a row not
// marked copies its decisions forward, a marked row clears its mark and
// recomputes.  Rows whose bits change mark their targets in dirty_out for
// the next round.  dirty_in NULL: every row recomputes (the first round).
__global__ void __launch_bounds__(256)
    k_fin_round(int n, const unsigned char *__restrict__ cls, const int *__restrict__ in_start,
                const int *__restrict__ in_sorted, const unsigned *__restrict__ dec_prev,
                unsigned *__restrict__ dec_next, const int *__restrict__ chg_prev,
                int *__restrict__ chg_cur, int *__restrict__ failed,
                const int32_t *__restrict__ face_ids, int *__restrict__ dirty_in,
                int *__restrict__ dirty_out) {
  const int sub = (int)(threadIdx.x & 7);
  const int i = (int)(blockIdx.x * 32 + (threadIdx.x >> 3));
  const bool live = i < n;
  // Every load that does not depend on another is issued before the first
  // branch, so the round costs one memory latency before its gathers (and
  // a skipped row none after it).
  const int go = chg_prev ? *chg_prev : 1;
  const int mark = (dirty_in && live) ? dirty_in[i] : 1;
  const unsigned dprev = live ? dec_prev[i] : 0u;
  int b = 0, e = 0;
  uint4 lo = make_uint4(0u, 0u, 0u, 0u), hi = lo;
  // this lane's four slots of the row's partners (slots 4*sub .. 4*sub+3),
  // for marking the targets of changed bits
  int4 tgt = make_int4(-1, -1, -1, -1);
  if (live && dirty_out)
    tgt = ((const int4 *)(face_ids + (int64_t)i * FOS_MAX_NEIGHBORS))[sub];
  if (live) {
    b = in_start[i];
    e = in_start[i + 1];
    if (sub == 0) {
      const uint4 *c4 = (const uint4 *)(cls + (int64_t)i * FOS_MAX_NEIGHBORS);
      lo = c4[0];
      hi = c4[1];
    }
  }
  if (go == 0) return;
  // the pair's first round changed something
  if (failed && blockIdx.x == 0 && threadIdx.x == 0 && *chg_cur) atomicOr(failed, 1);
  if (dirty_in && live) {
    // the 8 lanes of the row read the same mark, so they leave together
    if (mark == 0) {
      if (sub == 0) dec_next[i] = dprev;
      return;
    }
    if (sub == 0) dirty_in[i] = 0;
  }
  int acc = 0;
  if (live) {
    for (int p = b + sub; p < e; p += 8) {
      const int src = in_sorted[p];
      acc += (dec_prev[src / FOS_MAX_NEIGHBORS] >> (src % FOS_MAX_NEIGHBORS)) & 1u;
    }
  }
  acc += __shfl_xor(acc, 1, 8);
  acc += __shfl_xor(acc, 2, 8);
  acc += __shfl_xor(acc, 4, 8);
  unsigned bits = 0;
  if (live && sub == 0) {
    int used = acc < FOS_MAX_NEIGHBORS ? acc : FOS_MAX_NEIGHBORS;
    const unsigned words[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
    for (int s = 0; s < FOS_MAX_NEIGHBORS; ++s) {
      const unsigned c = (words[s >> 2] >> (8 * (s & 3))) & 0xffu;
      const unsigned k = c & 3u;
      const bool a = (k == CLS_ACCEPT) || (k == CLS_AMBIG && ((s < used) || (c & CLS_OLD)));
      if (a) {
        bits |= 1u << s;
        used += used < FOS_MAX_NEIGHBORS ? 1 : 0;
      }
    }
    dec_next[i] = bits;
  }
  // the row's decisions to all 8 lanes, which mark the targets of their
  // own changed slots
  bits = (unsigned)__shfl((int)bits, 0, 8);
  if (!live) return;
  const unsigned diff = bits ^ dprev;
  // one flag update per wave, not one per changed row: atomics on one
  // address serialise (round 0 changes nearly every row)
  const unsigned long long any = __ballot(diff != 0u);
  if (any && (int)(threadIdx.x & 63) == __ffsll((long long)any) - 1) {
    atomicOr(chg_cur, 1);
    if (failed) atomicOr(failed, 1);
  }
  if (diff == 0) return;
  if (dirty_out) {
    const unsigned mine = (diff >> (4 * sub)) & 0xfu;
    const int t4[4] = {tgt.x, tgt.y, tgt.z, tgt.w};
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (((mine >> q) & 1u) && t4[q] >= 0 && t4[q] < n) dirty_out[t4[q]] = 1;
  }
}

// Clears what finalize accumulates into: incoming counts (n + 1), fill
// cursors, decisions and both dirty-row mark arrays (n) and the per-pair
// change flags (nchg).
__global__ void k_fin_clear(int n, int *__restrict__ in_count, int *__restrict__ in_fill,
                            unsigned *__restrict__ dec, int *__restrict__ chg, int nchg,
                            int *__restrict__ dirty0, int *__restrict__ dirty1) {
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t <= n) in_count[t] = 0;
  if (t < n) {
    in_fill[t] = 0;
    dec[t] = 0;
    dirty0[t] = 0;
    dirty1[t] = 0;
  }
  if (t < nchg) chg[t] = 0;
}

// Writes row i's final neighbour list, hysteresis scores and counts.  The
// stores to prev follow the reference's chronology -- appends from lower
// rows first, then row i's own slots in order (an accept appends at `used`,
// a reject zeroes its own slot) -- and a later store to the same address
// wins, so prev ends exactly as in the serial loop; untouched slots keep last
// frame's value.
// 32 lanes per row (8 rows per 256-thread block), lane s owning position s.
// The serial chronology is resolved without racing stores: the appends from
// lower rows take positions 0 .. u_in - 1 (u_in = min(32, accepted
// incoming)); an accepted own slot s appends at u_in + (accepted slots
// before s), time s; a rejected own slot s zeroes position s, time s.  The
// store with the latest time is the one that remains.
__global__ void __launch_bounds__(256)
    k_fin_materialize(int n, const int32_t *__restrict__ face_ids,
                      const float *__restrict__ score, const unsigned char *__restrict__ cls,
                      const unsigned *__restrict__ dec, const int *__restrict__ in_start,
                      const int *__restrict__ in_sorted, float *__restrict__ prev,
                      int32_t *__restrict__ nl, int32_t *__restrict__ nc,
                      int32_t *__restrict__ used_count, int32_t *__restrict__ fsc,
                      float *__restrict__ ema, float alpha, int *__restrict__ zero_cnt,
                      const int *__restrict__ hold_a, const int *__restrict__ hold_b) {
  // a deferred face step that the host must still finish: the whole block
  // leaves before any barrier
  if ((hold_a && *hold_a) || (hold_b && *hold_b)) return;
  __shared__ int s_zero;
  __shared__ int s_time[8][FOS_MAX_NEIGHBORS];
  __shared__ float s_val[8][FOS_MAX_NEIGHBORS];
  __shared__ int s_id[8][FOS_MAX_NEIGHBORS];
  const int r = (int)(threadIdx.x >> 5), s = (int)(threadIdx.x & 31);
  const int half = (int)(threadIdx.x & 32);
  const unsigned below = (1u << s) - 1u;
  const int i = (int)blockIdx.x * 8 + r;
  const bool live = i < n;
  const int64_t base = (int64_t)(live ? i : 0) * FOS_MAX_NEIGHBORS;
  s_time[r][s] = -1;
  if (threadIdx.x == 0) s_zero = 0;
  __syncthreads();
  // appends from lower rows, in incoming order
  int u = 0;
  if (live) {
    const int e = in_start[i + 1];
    for (int p0 = in_start[i]; p0 < e && u < FOS_MAX_NEIGHBORS; p0 += FOS_MAX_NEIGHBORS) {
      const int p = p0 + s;
      int src = 0;
      bool acc = false;
      if (p < e) {
        src = in_sorted[p];
        acc = (dec[src / FOS_MAX_NEIGHBORS] >> (src % FOS_MAX_NEIGHBORS)) & 1u;
      }
      const unsigned m = (unsigned)(__ballot(acc) >> half);
      const int pos = u + __popc(m & below);
      if (acc && pos < FOS_MAX_NEIGHBORS) {
        s_val[r][pos] = score[src];
        s_id[r][pos] = src / FOS_MAX_NEIGHBORS;
      }
      u += __popc(m);
    }
  }
  const int u_in = u < FOS_MAX_NEIGHBORS ? u : FOS_MAX_NEIGHBORS;
  // the row's own slots
  const unsigned c = live ? (unsigned)cls[base + s] : (unsigned)CLS_INVALID;
  const bool valid = (c & 3u) != CLS_INVALID;
  const bool accd = valid && ((dec[live ? i : 0] >> s) & 1u);
  const bool rej = valid && !accd;
  const unsigned am = (unsigned)(__ballot(accd) >> half);
  const int rank = u_in + __popc(am & below);
  if (accd && rank < FOS_MAX_NEIGHBORS) {
    s_time[r][rank] = s;
    s_val[r][rank] = score[base + s];
    s_id[r][rank] = face_ids[base + s];
  }
  __syncthreads();
  if (live) {
    const int total = u_in + __popc(am);
    const int used = total < FOS_MAX_NEIGHBORS ? total : FOS_MAX_NEIGHBORS;
    const int ta = s_time[r][s], tr = rej ? s : -1;
    if (ta >= 0 || tr >= 0)
      prev[base + s] = ta > tr ? s_val[r][s] : 0.0f;
    else if (s < u_in)
      prev[base + s] = s_val[r][s];
    nl[base + s] = s < used ? s_id[r][s] : -1;
    if (s == 0) {
      nc[i] = used;
      used_count[i] = used;
      fsc[i] = used;
      // jfa.py:766-785 update_fsc_ema on the row's new FSC
      if (ema) {
        const float e = ema[i];
        ema[i] = (e == 0.0f && used > 0) ? (float)used : alpha * (float)used + (1.0f - alpha) * e;
      }
      if (zero_cnt && used == 0) atomicAdd(&s_zero, 1);
    }
  }
  if (zero_cnt) {
    __syncthreads();
    if (threadIdx.x == 0 && s_zero) atomicAdd(zero_cnt, s_zero);
  }
}

// jfa.py:766-785
__global__ void k_fsc_ema(int n, float alpha, const int32_t *__restrict__ fsc,
                          float *__restrict__ ema) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float e = ema[i];
  const int f = fsc[i];
  if (e == 0.0f && f > 0)
    ema[i] = (float)f;
  else
    ema[i] = alpha * (float)f + (1.0f - alpha) * e;
}

// ---------------------------------------------------------------------------
// Telemetry (jfa.py:1204-1330) and validation (jfa.py:1337-1397)
// ---------------------------------------------------------------------------

// acc: [0] p_dir, [1] overflow, [2] min, [3] max, [4] sum_fsc (int64 as
// unsigned long long); dacc: [0] sum_ema, [1] sum_score, [2] count_score
__global__ void k_stats_a(int n, const float *__restrict__ rad, float voxel,
                          const int32_t *__restrict__ fsc,
                          const float *__restrict__ ema,
                          const int32_t *__restrict__ used,
                          const int32_t *__restrict__ face_ids,
                          const int32_t *__restrict__ face_counts,
                          unsigned long long *acc, int *mm, double *dacc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int u = used[i];
  atomicAdd(&acc[0], (unsigned long long)u);
  if (u >= FOS_MAX_NEIGHBORS) atomicAdd(&acc[1], 1ull);
  atomicAdd(&acc[4], (unsigned long long)(long long)fsc[i]);
  atomicMin(&mm[0], fsc[i]);
  atomicMax(&mm[1], fsc[i]);
  atomicAdd(&dacc[0], (double)ema[i]);
  double s = 0.0;
  int c = 0;
  for (int k = 0; k < u && k < FOS_MAX_NEIGHBORS; ++k) {
    const int j = face_ids[(int64_t)i * FOS_MAX_NEIGHBORS + k];
    if (j < 0 || j >= n) continue;
    const int v = face_counts[(int64_t)i * FOS_MAX_NEIGHBORS + k];
    if (v <= 0) continue;
    const float e = expected_area(rad[i], rad[j], voxel);
    if (e > 0.0f) {
      s += (double)((float)v / e);
      c += 1;
    }
  }
  if (c) {
    atomicAdd(&dacc[1], s);
    atomicAdd(&dacc[2], (double)c);
  }
}

__global__ void k_stats_b(int n, const float *__restrict__ rad, float voxel,
                          const int32_t *__restrict__ used,
                          const int32_t *__restrict__ face_ids,
                          const int32_t *__restrict__ face_counts, double *dacc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double mean = dacc[2] > 0 ? dacc[1] / dacc[2] : 0.0;
  const int u = used[i];
  double s = 0.0;
  for (int k = 0; k < u && k < FOS_MAX_NEIGHBORS; ++k) {
    const int j = face_ids[(int64_t)i * FOS_MAX_NEIGHBORS + k];
    if (j < 0 || j >= n) continue;
    const int v = face_counts[(int64_t)i * FOS_MAX_NEIGHBORS + k];
    if (v <= 0) continue;
    const float e = expected_area(rad[i], rad[j], voxel);
    if (e > 0.0f) {
      const double d = (double)((float)v / e) - mean;
      s += d * d;
    }
  }
  if (s != 0.0) atomicAdd(&dacc[3], s);
}

__global__ void k_validate(int n, const int32_t *__restrict__ nl,
                           const int32_t *__restrict__ nc,
                           unsigned long long *out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * FOS_MAX_NEIGHBORS) return;
  const int i = (int)(t / FOS_MAX_NEIGHBORS), s = (int)(t % FOS_MAX_NEIGHBORS);
  const int j = nl[t];
  if (j == i) atomicAdd(&out[2], 1ull);
  if (j >= 0 && j < n) {
    bool found = false;
    for (int k = 0; k < FOS_MAX_NEIGHBORS; ++k)
      if (nl[(int64_t)j * FOS_MAX_NEIGHBORS + k] == i) {
        found = true;
        break;
      }
    if (!found) atomicAdd(&out[0], 1ull);
  }
  if (s == 0 && nc[i] >= FOS_MAX_NEIGHBORS) atomicAdd(&out[1], 1ull);
}

// jfa.py:983-1008; voxel -> tile (jfa.py:316-334) and wrapped halo marking
__global__ void k_dirty_mark(const float *__restrict__ pos,
                             const float *__restrict__ rad,
                             const float *__restrict__ pos_prev,
                             const float *__restrict__ rad_prev, int n, int res,
                             float L, float half_l, float voxel, int tile,
                             int halo, float pos_thr, float rad_thr,
                             uint8_t *__restrict__ dirty) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float dx = pos[3 * i] - pos_prev[3 * i], dy = pos[3 * i + 1] - pos_prev[3 * i + 1],
              dz = pos[3 * i + 2] - pos_prev[3 * i + 2];
  float d2 = dx * dx + dy * dy;
  d2 = d2 + dz * dz;
  const float pd = sqrtf(d2);
  const float rd = fabsf(rad[i] - rad_prev[i]);
  int tc[3], tp[3];
  for (int d = 0; d < 3; ++d) {
    tc[d] = wrap_fast((int)floorf(((pos[3 * i + d] + half_l) / L) * (float)res), res) / tile;
    tp[d] = wrap_fast((int)floorf(((pos_prev[3 * i + d] + half_l) / L) * (float)res), res) / tile;
  }
  const bool is_dirty = (pd > pos_thr * voxel) || (rd > rad_thr * voxel) ||
                        tc[0] != tp[0] || tc[1] != tp[1] || tc[2] != tp[2];
  if (!is_dirty) return;
  const int tpa = (res + tile - 1) / tile;
  for (int a = -halo; a <= halo; ++a)
    for (int b = -halo; b <= halo; ++b)
      for (int c = -halo; c <= halo; ++c) {
        const int t0 = wrap_fast(tc[0] + c, tpa), t1 = wrap_fast(tc[1] + b, tpa),
                  t2 = wrap_fast(tc[2] + a, tpa);
        dirty[((int64_t)t0 * tpa + t1) * tpa + t2] = 1;
      }
}

__global__ void k_dirty_count(const uint8_t *__restrict__ d, int64_t m,
                              unsigned long long *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int v = (i < m && d[i] == 1) ? 1 : 0;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(out, (unsigned long long)v);
}

__global__ void k_tile_cache(const float *__restrict__ pos,
                             const float *__restrict__ rad, int n,
                             float *__restrict__ pp, float *__restrict__ rp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  pp[3 * i] = pos[3 * i];
  pp[3 * i + 1] = pos[3 * i + 1];
  pp[3 * i + 2] = pos[3 * i + 2];
  rp[i] = rad[i];
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

static inline dim3 vox_grid(int res) {
  return dim3((unsigned)((res + 63) / 64), (unsigned)((res + 3) / 4), (unsigned)res);
}

static inline bool unit_l(float L) { return L == 1.0f; }

extern "C" {

int fos_jfa_seed_table(fos_ctx *ctx, const float *pos, const float *rad, int n,
                       float beta, void *stream) {
  FOS_REQUIRE(ctx && n >= 0 && n <= ctx->max_n && (n == 0 || (pos && rad)),
              "fos_jfa_seed_table: bad arguments (n=%d)", n);
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_seed_table, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream, pos, rad, n, beta, ctx->seeds);
  FOS_CHECK_LAUNCH("k_seed_table");
  return 0;
}

int fos_jfa_init(fos_ctx *ctx, const float *pos, const float *rad, int n,
                 int res, float L, float voxel, float beta, int32_t *grid,
                 int32_t *temp, void *stream) {
  FOS_REQUIRE(ctx && grid && res >= 1 && res <= ctx->max_res && n >= 0 &&
                  n <= ctx->max_n && L > 0.0f && voxel > 0.0f,
              "fos_jfa_init: bad arguments (n=%d res=%d)", n, res);
  hipStream_t st = (hipStream_t)stream;
  const int64_t V = (int64_t)res * res * res;
  int rc = fos_fill_i32(grid, V, -1, stream);
  if (rc) return rc;
  if (temp) {
    rc = fos_fill_i32(temp, V, -1, stream);
    if (rc) return rc;
  }
  if (n == 0) return 0;
  FOS_REQUIRE(pos && rad, "fos_jfa_init: NULL pos or rad");
  const float half_l = (float)(0.5 * (double)L);
  // Pass slots 0 .. FOS_PASS_SLOTS - 2 (what fos_jfa_passes uses) are
  // cleared here too; fos_jfa_passes skips its own clear while that holds.
  const int ncl = (FOS_PASS_SLOTS - 1) * FOS_STRIPE_PASS_INTS;
  if (unit_l(L))
    hipLaunchKernelGGL(k_seed_fused<true>, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, st, pos,
                       rad, beta, ctx->seeds, n, res, L, half_l, voxel, grid, ctx->stripes, ncl);
  else
    hipLaunchKernelGGL(k_seed_fused<false>, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, st, pos,
                       rad, beta, ctx->seeds, n, res, L, half_l, voxel, grid, ctx->stripes, ncl);
  FOS_CHECK_LAUNCH("k_seed_fused");
  ctx->stripes_clear = 1;
  return 0;
}

}  // extern "C"

static int g_emit_tx = 8;
static int g_emit_zw = 32;
// fos_set_param "fin_first_burst": rounds of the finalize's first burst
// (0, default: what the last finalize needed, 2 to 16)
static int g_fin_first = 0;
extern int g_grid_block;  // fos_grid.hip
extern int g_grid_cpw;    // fos_grid.hip

// @jfa-pass-launch-begin
// fos_set_param "jfa_variant": bit 0 evaluates four list entries per
// iteration (two otherwise), bit 1 carries each voxel's winning power
// distance into the next pass (the incumbent is not evaluated again), bit 2
// takes the x-planes in residue-class order; -1 (default) picks four
// entries for 32-bit lists (ids above 65,535: the large grids, where it was
// measured faster), the distance carry whenever the context holds the
// distance buffers (grids up to 256^3) and the class order per
// jfa_class_order.
static int g_jfa_variant = -1;
// fos_set_param "jfa_half_orbits": 1 runs the step = res / 2 pass by orbits
// (k_jfa_half), 0 with the list kernel's JFA_HALF flavour, -1 (default) by
// orbits on grids of 256^3 and more.  One thread per orbit loads each label
// once where the list kernel loads it eight times, which pays where the
// labels come from HBM (C3: 418 -> 167 us); on a grid that fits in the L2s
// the list kernel's eight times as many waves hide latency better (C2: 18.9
// against 23.2 us).
static int g_jfa_half_orbits = -1;
static int jfa_variant(const fos_ctx *ctx) {
  if (g_jfa_variant >= 0) return g_jfa_variant;
  return (ctx->max_n > 65536 ? 1 : 0) | (ctx->dist[0] ? 2 : 0);
}

// Whether a pass takes its x-planes in residue-class order (jfa_plane).  By
// default only where it was measured faster: grids of 256^3 and more, whose
// x-slab per XCD exceeds its L2, at steps 2 .. res/16 (at larger steps the
// classes are short and the planes they share are few; at step 1 the order
// is the natural one).
static bool jfa_class_order(const fos_ctx *ctx, int res, int step) {
  if (g_jfa_variant >= 0) return (g_jfa_variant & 4) != 0;
  return res >= 256 && step >= 2 && step * 16 <= res;
}

template <bool U, int MODE, typename LT, int BATCH, int DM>
static void launch_pass_list_v(fos_ctx *ctx, int res, float L, float half_l, float voxel,
                               int step, const int32_t *in, int32_t *out, int pass, int early,
                               double th, int minp, const float *dist_in, float *dist_out,
                               hipStream_t st) {
  const int nyb = (res + 3) / 4, nzb = (res + 63) / 64;
  const int ntiles = res * nyb * nzb;
  const int per = (ntiles + FOS_XCDS - 1) / FOS_XCDS;
  const unsigned q = (unsigned)(res / step);
  const int qr = (int)((q << 16) | (unsigned)(res - (int)q * step));
  const unsigned mq = (unsigned)(((1ull << 32) + q - 1) / q);
  const unsigned mq1 = (unsigned)(((1ull << 32) + q) / (q + 1));
  if (jfa_class_order(ctx, res, step))
    hipLaunchKernelGGL((k_jfa_pass_list<U, MODE, LT, BATCH, DM, true>), dim3(per * FOS_XCDS),
                       dim3(256), 0, st, ctx->seeds, res, L, half_l, voxel, step, in, out,
                       ctx->stripes, pass, early, th, minp, nyb, nzb, ntiles, per, dist_in,
                       dist_out, qr, mq, mq1);
  else
    hipLaunchKernelGGL((k_jfa_pass_list<U, MODE, LT, BATCH, DM>), dim3(per * FOS_XCDS),
                       dim3(256), 0, st, ctx->seeds, res, L, half_l, voxel, step, in, out,
                       ctx->stripes, pass, early, th, minp, nyb, nzb, ntiles, per, dist_in,
                       dist_out, qr, mq, mq1);
}

// dm: 0 no distance buffers, 1 write dist_out, 2 read dist_in and write dist_out
template <bool U, int MODE>
static void launch_pass_list(fos_ctx *ctx, int res, float L, float half_l, float voxel,
                             int step, const int32_t *in, int32_t *out, int pass, int early,
                             double th, int minp, int dm, const float *dist_in, float *dist_out,
                             hipStream_t st) {
  const bool b4 = (jfa_variant(ctx) & 1) != 0;
#define FOS_LV(LT, B, D)                                                                     \
  launch_pass_list_v<U, MODE, LT, B, D>(ctx, res, L, half_l, voxel, step, in, out, pass,     \
                                        early, th, minp, dist_in, dist_out, st)
#define FOS_LD(LT, B)                                                                        \
  do {                                                                                       \
    if (dm == 2) FOS_LV(LT, B, 2);                                                           \
    else if (dm == 1) FOS_LV(LT, B, 1);                                                      \
    else FOS_LV(LT, B, 0);                                                                   \
  } while (0)
  // owner ids are below the context's particle capacity: 16-bit list
  // entries when that is at most 65536
  if (ctx->max_n <= 65536) {
    if (b4) FOS_LD(unsigned short, 4); else FOS_LD(unsigned short, 2);
  } else {
    if (b4) FOS_LD(int, 4); else FOS_LD(int, 2);
  }
#undef FOS_LD
#undef FOS_LV
}

// Launches pass `pass` (its changed-voxel count accumulates in stripe slot
// `pass`; the caller zeroes the slot first).  `settled` marks a pass whose
// step equals the previous pass's (the field is close to converged): the
// list kernel then also skips owners repeated by -1 neighbours.  dm and the
// distance buffers as for launch_pass_list.
static int launch_step(fos_ctx *ctx, int res, float L, float voxel, int step,
                       const int32_t *in, int32_t *out, int pass, int early, double th,
                       int min_passes, bool settled, hipStream_t st, int dm = 0,
                       const float *dist_in = nullptr, float *dist_out = nullptr) {
  const float half_l = (float)(0.5 * (double)L);
  const bool u = unit_l(L);
  const bool orbits = g_jfa_half_orbits < 0 ? res >= 256 : g_jfa_half_orbits != 0;
  if (2 * step == res && orbits && dm != 2) {
    // the step = res / 2 pass by orbits (k_jfa_half)
    const int s = res / 2;
    const unsigned b = blocks_for((int64_t)s * s * s);
#define FOS_HALF(U, D)                                                                           \
  hipLaunchKernelGGL((k_jfa_half<U, D>), dim3(b), dim3(256), 0, st, ctx->seeds, res, L, half_l, \
                     voxel, in, out, ctx->stripes, pass, early, th, min_passes, dist_out)
    if (u) {
      if (dm) FOS_HALF(true, 1); else FOS_HALF(true, 0);
    } else {
      if (dm) FOS_HALF(false, 1); else FOS_HALF(false, 0);
    }
#undef FOS_HALF
    FOS_CHECK_LAUNCH("k_jfa_half");
    return 0;
  }
  const int mode = 2 * step == res ? JFA_HALF : settled ? JFA_SETTLED : JFA_PLAIN;
#define FOS_LIST(U, M)                                                                   \
  launch_pass_list<U, M>(ctx, res, L, half_l, voxel, step, in, out, pass, early, th,    \
                         min_passes, dm, dist_in, dist_out, st)
  if (u) {
    if (mode == JFA_HALF) FOS_LIST(true, JFA_HALF);
    else if (mode == JFA_SETTLED) FOS_LIST(true, JFA_SETTLED);
    else FOS_LIST(true, JFA_PLAIN);
  } else {
    if (mode == JFA_HALF) FOS_LIST(false, JFA_HALF);
    else if (mode == JFA_SETTLED) FOS_LIST(false, JFA_SETTLED);
    else FOS_LIST(false, JFA_PLAIN);
  }
#undef FOS_LIST
  FOS_CHECK_LAUNCH("k_jfa_pass_list");
  return 0;
}

// @jfa-pass-launch-end

// Face-pair emission into the row tables (the kernel chosen by
// face_emit_tx); the overflow flag (ctx->flags[0]) is left for the caller.
static int face_emit(fos_ctx *ctx, int res, int n, const int32_t *filtered, hipStream_t st) {
  const int tx = g_emit_tx;
  if (tx == 0) {
    const int nyb = (res + 3) / 4, nzb = (res + 63) / 64, ntiles = res * nyb * nzb;
    const int per = (ntiles + FOS_XCDS - 1) / FOS_XCDS;
    hipLaunchKernelGGL(k_face_emit, dim3(per * FOS_XCDS), dim3(256), 0, st, res, n, filtered,
                       ctx->tab_b, ctx->tab_k, ctx->tab_c, ctx->tab_cap, ctx->flags, nyb, nzb,
                       ntiles, per);
  } else {
    const int zw = g_emit_zw, yt = 256 / zw;
    const int nyb = (res + yt - 1) / yt, nzb = (res + zw - 1) / zw;
    const int ntiles_x = ((res + tx - 1) / tx) * nyb * nzb;
    const int per_x = (ntiles_x + FOS_XCDS - 1) / FOS_XCDS;
#define EMIT_K(T)                                                             \
  (zw == 64 ? k_face_emit_lds<T, 64> : zw == 32 ? k_face_emit_lds<T, 32> : \
              k_face_emit_lds<T, 16>)
    auto kern = tx == 1   ? EMIT_K(1)
                : tx == 2 ? EMIT_K(2)
                : tx == 4 ? EMIT_K(4)
                : tx == 8 ? EMIT_K(8)
                          : EMIT_K(16);
#undef EMIT_K
    hipLaunchKernelGGL(kern, dim3(per_x * FOS_XCDS), dim3(256), 0, st, res, n, filtered,
                       ctx->tab_b, ctx->tab_k, ctx->tab_c, ctx->tab_cap, ctx->flags, nyb, nzb,
                       ntiles_x, per_x);
  }
  FOS_CHECK_LAUNCH("k_face_emit");
  return 0;
}

// After an overflow: tables four times larger (cleared), flag cleared.
static int face_grow(fos_ctx *ctx, hipStream_t st) {
  const int cap = ctx->tab_cap * 4;
  FOS_REQUIRE(cap <= 4096, "face accumulation: row table overflow at %d", ctx->tab_cap);
  int rc = alloc_tables(ctx, cap);
  if (rc) return rc;
  return fos_fill_i32(ctx->flags, 1, 0, st);
}

static int *fin_chg(fos_ctx *ctx);
static int *fin_dirty(fos_ctx *ctx, int which);
#define FIN_RING 48

// clear_fin: also clear the finalize accumulators (a finalize follows).
static int face_compact(fos_ctx *ctx, int n, int32_t *face_ids, int32_t *face_counts,
                        hipStream_t st, bool clear_fin = false, int *zero_cnt = nullptr) {
  const int cap = ctx->tab_cap;
  const size_t shmem = (size_t)4 * cap * sizeof(unsigned);
  hipLaunchKernelGGL(k_face_compact, dim3((unsigned)((n + 3) / 4)), dim3(256), shmem, st, n,
                     ctx->tab_b, ctx->tab_k, ctx->tab_c, cap, face_ids, face_counts,
                     clear_fin ? ctx->in_count : nullptr, ctx->in_fill, (unsigned *)ctx->dec,
                     fin_chg(ctx), FIN_RING, fin_dirty(ctx, 0), fin_dirty(ctx, 1), zero_cnt);
  FOS_CHECK_LAUNCH("k_face_compact");
  return 0;
}

extern "C" {

int fos_set_param(const char *name, int value) {
  FOS_REQUIRE(name, "fos_set_param: name is NULL");
  if (!strcmp(name, "jfa_half_orbits")) {
    FOS_REQUIRE(value >= -1 && value <= 1, "jfa_half_orbits must be -1, 0 or 1");
    g_jfa_half_orbits = value;
  } else if (!strcmp(name, "jfa_variant")) {
    FOS_REQUIRE(value >= -1 && value <= 7, "jfa_variant must be in [-1, 7]");
    g_jfa_variant = value;
  } else if (!strcmp(name, "face_emit_tx")) {
    FOS_REQUIRE(value == 0 || value == 1 || value == 2 || value == 4 || value == 8 ||
                    value == 16,
                "face_emit_tx must be 0, 1, 2, 4, 8 or 16");
    g_emit_tx = value;
  } else if (!strcmp(name, "face_emit_zw")) {
    FOS_REQUIRE(value == 16 || value == 32 || value == 64, "face_emit_zw must be 16, 32 or 64");
    g_emit_zw = value;
  } else if (!strcmp(name, "grid_block")) {
    FOS_REQUIRE(value == 0 || value == 1, "grid_block must be 0 or 1");
    g_grid_block = value;
  } else if (!strcmp(name, "pbd_lanes")) {
    return pbd_lanes_set(value);
  } else if (!strcmp(name, "pressure_flow")) {
    return pressure_flow_set(value);
  } else if (!strcmp(name, "fin_first_burst")) {
    FOS_REQUIRE(value == 0 || (value >= 2 && value <= 16), "fin_first_burst must be 0 or in [2, 16]");
    g_fin_first = value;
  } else if (!strcmp(name, "grid_range_cells")) {
    FOS_REQUIRE(value == 0 || (value >= 1 && value <= FOS_GR_CELLS),
                "grid_range_cells must be 0 or in [1, %d]", FOS_GR_CELLS);
    g_grid_cpw = value;
  } else {
    FOS_REQUIRE(false, "fos_set_param: unknown parameter '%s'", name);
  }
  return 0;
}

int fos_get_param(const char *name, int *value) {
  FOS_REQUIRE(name && value, "fos_get_param: bad arguments");
  if (!strcmp(name, "jfa_variant"))
    *value = g_jfa_variant;
  else if (!strcmp(name, "jfa_half_orbits"))
    *value = g_jfa_half_orbits;
  else if (!strcmp(name, "grid_block"))
    *value = g_grid_block;
  else if (!strcmp(name, "face_emit_tx"))
    *value = g_emit_tx;
  else if (!strcmp(name, "face_emit_zw"))
    *value = g_emit_zw;
  else if (!strcmp(name, "pbd_lanes"))
    *value = pbd_lanes_get();
  else if (!strcmp(name, "pressure_flow"))
    *value = pressure_flow_get();
  else if (!strcmp(name, "fin_first_burst"))
    *value = g_fin_first;
  else if (!strcmp(name, "grid_range_cells"))
    *value = g_grid_cpw;
  else
    FOS_REQUIRE(false, "fos_get_param: unknown parameter '%s'", name);
  return 0;
}

int fos_jfa_step(fos_ctx *ctx, int res, float L, float voxel, int step_size,
                 const int32_t *in, int32_t *out, int32_t *changes,
                 void *stream) {
  FOS_REQUIRE(ctx && in && out && in != out && res >= 1 && res <= ctx->max_res &&
                  step_size >= 1 && step_size <= res && L > 0.0f,
              "fos_jfa_step: bad arguments (res=%d step=%d)", res, step_size);
  hipStream_t st = (hipStream_t)stream;
  const int slot = FOS_PASS_SLOTS - 1;
  int *sl = ctx->stripes + slot * FOS_STRIPE_PASS_INTS;
  int rc = fos_fill_i32(sl, FOS_STRIPE_PASS_INTS, 0, stream);
  if (rc) return rc;
  rc = launch_step(ctx, res, L, voxel, step_size, in, out, slot, 0, 0.0, 0, step_size == 1, st);
  if (rc || !changes) return rc;
  hipLaunchKernelGGL(k_stripes_add, dim3(1), dim3(64), 0, st, (const int *)ctx->stripes, slot,
                     changes);
  FOS_CHECK_LAUNCH("k_stripes_add");
  return 0;
}

int fos_jfa_copy(int res, const int32_t *src, int32_t *dst, void *stream) {
  FOS_REQUIRE(src && dst && res >= 1, "fos_jfa_copy: bad arguments");
  const int64_t V = (int64_t)res * res * res;
  unsigned b = blocks_for(V);
  if (b > 8192) b = 8192;
  hipLaunchKernelGGL(k_copy_i32, dim3(b), dim3(FOS_BLOCK), 0, (hipStream_t)stream,
                     src, dst, V);
  FOS_CHECK_LAUNCH("k_copy_i32");
  return 0;
}

int fos_jfa_passes_timed(fos_ctx *ctx, int res, float L, float voxel,
                         int num_passes, int early_exit, double th_pass_delta,
                         int min_passes, int32_t *grid, int32_t *temp,
                         int32_t *info, void *const *events, void *stream) {
  FOS_REQUIRE(ctx && grid && temp && info && grid != temp && res >= 1 &&
                  res <= ctx->max_res && num_passes >= 0 && num_passes <= 64 &&
                  L > 0.0f,
              "fos_jfa_passes: bad arguments (res=%d passes=%d)", res, num_passes);
  FOS_REQUIRE(num_passes < FOS_PASS_SLOTS, "fos_jfa_passes: at most %d passes",
              FOS_PASS_SLOTS - 1);
  hipStream_t st = (hipStream_t)stream;
  int rc = 0;
  if (!ctx->stripes_clear) {
    rc = fos_fill_i32(ctx->stripes, (int64_t)num_passes * FOS_STRIPE_PASS_INTS, 0, stream);
    if (rc) return rc;
  }
  ctx->stripes_clear = 0;
  // A pass is settled when its step equals the previous pass's.
  int step = res / 2, prev = 0;
  for (int p = 0; p < num_passes; ++p) {
    if (step < 1) step = 1;
    const int32_t *in = (p & 1) ? temp : grid;
    int32_t *out = (p & 1) ? grid : temp;
    if (events && events[p]) FOS_HIP(hipEventRecord((hipEvent_t)events[p], st));
    // winning power distances carried from pass to pass (the first pass
    // evaluates the seeded incumbents and writes them)
    const bool carry = ctx->dist[0] && (jfa_variant(ctx) & 2);
    rc = launch_step(ctx, res, L, voxel, step, in, out, p, early_exit, th_pass_delta,
                     min_passes, step == prev, st, carry ? (p ? 2 : 1) : 0,
                     carry ? ctx->dist[(p + 1) & 1] : nullptr,
                     carry ? ctx->dist[p & 1] : nullptr);
    if (rc) return rc;
    prev = step;
    step /= 2;
  }
  if (events && events[num_passes]) FOS_HIP(hipEventRecord((hipEvent_t)events[num_passes], st));
  const int64_t V = (int64_t)res * res * res;
  unsigned b = blocks_for(V);
  if (b > 2048) b = 2048;
  // without early exit every pass runs: an even count leaves the labels in
  // grid, and one block writes the totals
  if (!early_exit && num_passes % 2 == 0) b = 1;
  hipLaunchKernelGGL(k_jfa_finish, dim3(b), dim3(FOS_BLOCK), 0, st, res, num_passes,
                     early_exit, th_pass_delta, min_passes, (const int *)ctx->stripes, info,
                     temp, grid);
  FOS_CHECK_LAUNCH("k_jfa_finish");
  return 0;
}

int fos_jfa_passes(fos_ctx *ctx, int res, float L, float voxel, int num_passes,
                   int early_exit, double th_pass_delta, int min_passes,
                   int32_t *grid, int32_t *temp, int32_t *info, void *stream) {
  return fos_jfa_passes_timed(ctx, res, L, voxel, num_passes, early_exit, th_pass_delta,
                              min_passes, grid, temp, info, nullptr, stream);
}

int fos_events_create(int count, void **events) {
  FOS_REQUIRE(count >= 1 && events, "fos_events_create: bad arguments");
  for (int i = 0; i < count; ++i) {
    hipEvent_t e;
    FOS_HIP(hipEventCreate(&e));
    events[i] = (void *)e;
  }
  return 0;
}

int fos_events_destroy(int count, void **events) {
  FOS_REQUIRE(count >= 0 && (count == 0 || events), "fos_events_destroy: bad arguments");
  for (int i = 0; i < count; ++i)
    if (events[i]) (void)hipEventDestroy((hipEvent_t)events[i]);
  return 0;
}

int fos_events_elapsed(int count, void *const *events, float *ms) {
  FOS_REQUIRE(count >= 2 && events && ms, "fos_events_elapsed: bad arguments");
  FOS_HIP(hipEventSynchronize((hipEvent_t)events[count - 1]));
  for (int i = 0; i + 1 < count; ++i)
    FOS_HIP(hipEventElapsedTime(&ms[i], (hipEvent_t)events[i], (hipEvent_t)events[i + 1]));
  return 0;
}

int fos_majority(int res, const int32_t *grid, int32_t *filtered, void *stream) {
  FOS_REQUIRE(grid && filtered && grid != filtered && res >= 1,
              "fos_majority: bad arguments");
  const int nyb = (res + 3) / 4, nzb = (res + 63) / 64, ntiles = res * nyb * nzb;
  const int per = (ntiles + FOS_XCDS - 1) / FOS_XCDS;
  hipLaunchKernelGGL(k_majority, dim3(per * FOS_XCDS), dim3(256), 0, (hipStream_t)stream,
                     res, grid, filtered, nyb, nzb, ntiles, per);
  FOS_CHECK_LAUNCH("k_majority");
  return 0;
}

int fos_face_accumulate(fos_ctx *ctx, int res, int n, const int32_t *filtered,
                        int32_t *face_ids, int32_t *face_counts, void *stream) {
  FOS_REQUIRE(ctx && filtered && face_ids && face_counts && res >= 1 &&
                  res <= ctx->max_res && n >= 0 && n <= ctx->max_n,
              "fos_face_accumulate: bad arguments (n=%d res=%d)", n, res);
  FOS_REQUIRE((double)res * res * res * 6.0 < 4294967295.0,
              "fos_face_accumulate: res=%d exceeds the 32-bit face key", res);
  if (n == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  while (true) {
    int rc = face_emit(ctx, res, n, filtered, st);
    if (rc) return rc;
    FOS_HIP(hipMemcpyAsync(ctx->host_i, ctx->flags, sizeof(int), hipMemcpyDeviceToHost, st));
    FOS_HIP(hipStreamSynchronize(st));
    if (!*(volatile int *)ctx->host_i) break;
    rc = face_grow(ctx, st);
    if (rc) return rc;
  }
  return face_compact(ctx, n, face_ids, face_counts, st);
}


}  // extern "C"

static int finalize_impl(fos_ctx *ctx, int n, const float *rad, float voxel, const float *pvox,
                         const int32_t *face_ids, const int32_t *face_counts,
                         float *prev_scores, int32_t *neighbor_list, int32_t *neighbor_count,
                         int32_t *used_count, int32_t *fsc, int fixed_pairs, void *stream);

extern "C" {

int fos_finalize_faces(fos_ctx *ctx, int n, const float *rad, float voxel,
                       const int32_t *face_ids, const int32_t *face_counts,
                       float *prev_scores, int32_t *neighbor_list,
                       int32_t *neighbor_count, int32_t *used_count,
                       int32_t *fsc, void *stream) {
  FOS_REQUIRE(ctx && n >= 0 && n <= ctx->max_n && voxel > 0.0f &&
                  (n == 0 || (rad && face_ids && face_counts && prev_scores &&
                              neighbor_list && neighbor_count && used_count && fsc)),
              "fos_finalize_faces: bad arguments (n=%d)", n);
  return finalize_impl(ctx, n, rad, voxel, nullptr, face_ids, face_counts, prev_scores,
                       neighbor_list, neighbor_count, used_count, fsc, 0, stream);
}

}  // extern "C"

// Per-round change flags of the fixed point: a ring of FIN_RING at
// ctx->flags + 16.  FIN_BURST: rounds queued per host read after the first
// burst.
#define FIN_BURST 4
static int *fin_chg(fos_ctx *ctx) { return ctx->flags + 16; }

// Finalize in phases: prepare (classify, incoming lists, decisions cleared),
// round pairs, materialize.
// cleared: the accumulators were cleared already (face_compact with clear_fin).
static int fin_prepare(fos_ctx *ctx, int n, const float *rad, float voxel, const float *pvox,
                       const int32_t *face_ids, const int32_t *face_counts,
                       const float *prev_scores, hipStream_t st, bool cleared = false) {
  const int64_t m = (int64_t)n * FOS_MAX_NEIGHBORS;
  // the fixed point starts from "nothing accepted"
  if (!cleared) {
    hipLaunchKernelGGL(k_fin_clear,
                       dim3(blocks_for((int64_t)(n + 1 > FIN_RING ? n + 1 : FIN_RING))),
                       dim3(FOS_BLOCK), 0, st, n, ctx->in_count, ctx->in_fill,
                       (unsigned *)ctx->dec, fin_chg(ctx), FIN_RING, fin_dirty(ctx, 0),
                       fin_dirty(ctx, 1));
    FOS_CHECK_LAUNCH("k_fin_clear");
  }
  hipLaunchKernelGGL(k_fin_classify, dim3(blocks_for(m)), dim3(FOS_BLOCK), 0, st, n,
                     rad, voxel, pvox, face_ids, face_counts, prev_scores, ctx->score,
                     ctx->cls, ctx->in_count);
  FOS_CHECK_LAUNCH("k_fin_classify");
  int rc = exclusive_scan(ctx, ctx->in_count, ctx->in_start, n + 1, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_fin_scatter, dim3(blocks_for(m)), dim3(FOS_BLOCK), 0, st, n,
                     face_ids, ctx->cls, ctx->in_start, ctx->in_fill, ctx->in_src);
  FOS_CHECK_LAUNCH("k_fin_scatter");
  hipLaunchKernelGGL(k_fin_rank, dim3(blocks_for(m)), dim3(FOS_BLOCK), 0, st, n, face_ids,
                     ctx->in_start, ctx->in_src, ctx->in_sorted);
  FOS_CHECK_LAUNCH("k_fin_rank");
  return 0;
}

// Rounds r0 .. r0 + rounds - 1 with no host round trip.  Round r reads the
// decision buffer r & 1 (a, b, a, ...) and writes the other; its change
// flag is fin_chg(ctx)[r % FIN_RING], and a round after one that changed
// nothing returns at once.  Any fixed point of the round map is the serial
// answer (a row's decisions depend only on lower rows', so by induction over
// rows a fixed point agrees with the serial loop), so the first round that
// changes nothing ends the iteration, and it leaves both buffers equal.
// With `last_failed`, the last round marks ctx->flags[3] (sticky) when it
// changed anything.  Flags are cleared by fin_prepare; a caller going past
// FIN_RING rounds clears the ring itself (fin_ring_clear).
// Dirty-row marks of the rounds: two int arrays of max_n in the decision
// scratch after the two decision buffers; round r >= 2 reads (and clears)
// array r & 1 and marks array (r + 1) & 1.
static int *fin_dirty(fos_ctx *ctx, int which) {
  return (int *)ctx->dec + (size_t)(2 + which) * ctx->max_n;
}

static int fin_rounds(fos_ctx *ctx, int n, int r0, int rounds, bool last_failed,
                      const int32_t *face_ids, hipStream_t st) {
  unsigned *dec[2] = {(unsigned *)ctx->dec, (unsigned *)ctx->dec + ctx->max_n};
  int *chg = fin_chg(ctx);
  int *d[2] = {fin_dirty(ctx, 0), fin_dirty(ctx, 1)};
  for (int r = r0; r < r0 + rounds; ++r) {
    const int *prev = r > 0 ? chg + (r - 1) % FIN_RING : nullptr;
    int *cur = chg + r % FIN_RING;
    int *failed = last_failed && r == r0 + rounds - 1 ? ctx->flags + 3 : nullptr;
    // round 0 changes nearly every row (from nothing accepted), so it marks
    // nothing and round 1 recomputes every row
    hipLaunchKernelGGL(k_fin_round, dim3((unsigned)((n + 31) / 32)), dim3(256), 0, st, n,
                       ctx->cls, ctx->in_start, ctx->in_sorted, dec[r & 1], dec[(r + 1) & 1],
                       prev, cur, failed, face_ids, r >= 2 ? d[r & 1] : nullptr,
                       r >= 1 ? d[(r + 1) & 1] : nullptr);
    FOS_CHECK_LAUNCH("k_fin_round");
  }
  return 0;
}

// Clears every ring slot but `keep` (the flag of the pair before the next
// burst, which that burst's first pair reads).
static int fin_ring_clear(fos_ctx *ctx, int keep, hipStream_t st) {
  int rc = keep > 0 ? fos_fill_i32(fin_chg(ctx), keep, 0, st) : 0;
  if (!rc && keep + 1 < FIN_RING)
    rc = fos_fill_i32(fin_chg(ctx) + keep + 1, FIN_RING - keep - 1, 0, st);
  return rc;
}

// ema (may be NULL): also the FSC EMA update with alpha; zero_cnt (may be
// NULL, zero before): also counts the rows whose FSC is 0.
// hold_a / hold_b (may be NULL): device words; when either is set the
// kernel does nothing (a deferred face step the host must still finish).
static int fin_materialize(fos_ctx *ctx, int n, const int32_t *face_ids, float *prev_scores,
                           int32_t *neighbor_list, int32_t *neighbor_count, int32_t *used_count,
                           int32_t *fsc, hipStream_t st, float *ema = nullptr, float alpha = 0.0f,
                           int *zero_cnt = nullptr, const int *hold_a = nullptr,
                           const int *hold_b = nullptr) {
  hipLaunchKernelGGL(k_fin_materialize, dim3((unsigned)((n + 7) / 8)), dim3(256), 0, st, n,
                     face_ids, ctx->score, ctx->cls, (const unsigned *)ctx->dec, ctx->in_start,
                     ctx->in_sorted, prev_scores, neighbor_list, neighbor_count, used_count, fsc,
                     ema, alpha, zero_cnt, hold_a, hold_b);
  FOS_CHECK_LAUNCH("k_fin_materialize");
  return 0;
}

// Reads ctx->flags (all 64 ints: overflow, sticky failed, the pair ring)
// into host memory (one round trip).
static int read_flags(fos_ctx *ctx, int out[64], hipStream_t st) {
  FOS_HIP(hipMemcpyAsync(ctx->host_i, ctx->flags, 64 * sizeof(int), hipMemcpyDeviceToHost, st));
  FOS_HIP(hipStreamSynchronize(st));
  const volatile int *h = (const volatile int *)ctx->host_i;
  for (int k = 0; k < 64; ++k) out[k] = h[k];
  return 0;
}

// Rounds up to the first that changed nothing, from flags read after round
// `last` (the ring has not wrapped); FIN_BURST when it has.
static int fin_needed(const int *f, int last) {
  if (last >= FIN_RING) return FIN_BURST;
  for (int r = 0; r <= last; ++r)
    if (!f[16 + r]) return r + 1;
  return FIN_BURST;
}

// The first burst's size: what the last finalize needed (consecutive frames
// need about the same), 2 to 16 rounds.
static int fin_first_burst(const fos_ctx *ctx) {
  if (g_fin_first) return g_fin_first;
  const int b = ctx->fin_burst > 0 ? ctx->fin_burst : FIN_BURST;
  return b < 2 ? 2 : b > 16 ? 16 : b;
}

// Bursts of rounds from round r0 until the fixed point (a burst whose last
// round changed nothing), the host reading the flags after each burst: the
// first of `first` rounds, then FIN_BURST.  `known`: the flags of a burst
// already read (then r0 is the round after it); NULL starts with a burst.
// Records the rounds needed in ctx->fin_burst.
static int fin_rounds_host(fos_ctx *ctx, int n, int r0, int first, const int *known,
                           const int32_t *face_ids, hipStream_t st) {
  int f[64];
  if (known && !known[16 + (r0 - 1) % FIN_RING]) {
    ctx->fin_burst = fin_needed(known, r0 - 1);
    return 0;
  }
  // rounds below lap_end have ring slots not used since fin_prepare cleared them
  int lap_end = FIN_RING;
  for (int r = r0, p = first;; r += p, p = FIN_BURST) {
    FOS_REQUIRE(r <= 2 * n + 2, "finalize: no fixed point after %d rounds", r);
    int rc = 0;
    if (r + p > lap_end) {
      rc = fin_ring_clear(ctx, (r - 1) % FIN_RING, st);
      lap_end = r + FIN_RING - 1;
    }
    if (!rc) rc = fin_rounds(ctx, n, r, p, false, face_ids, st);
    if (!rc) rc = read_flags(ctx, f, st);
    if (rc) return rc;
    if (!f[16 + (r + p - 1) % FIN_RING]) {
      ctx->fin_burst = fin_needed(f, r + p - 1);
      return 0;
    }
  }
}

// fixed_pairs > 0: that many round pairs (2 rounds each) with no host round
// trip (flags[3] stays set if the last round had not reached the fixed
// point); 0: until the fixed point (fin_rounds_host).
static int finalize_impl(fos_ctx *ctx, int n, const float *rad, float voxel, const float *pvox,
                         const int32_t *face_ids, const int32_t *face_counts,
                         float *prev_scores, int32_t *neighbor_list, int32_t *neighbor_count,
                         int32_t *used_count, int32_t *fsc, int fixed_pairs, void *stream) {
  if (n == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  int rc = fin_prepare(ctx, n, rad, voxel, pvox, face_ids, face_counts, prev_scores, st);
  if (rc) return rc;
  if (fixed_pairs > 0) {
    FOS_REQUIRE(2 * fixed_pairs <= FIN_RING, "finalize: fixed_pairs %d above %d", fixed_pairs,
                FIN_RING / 2);
    rc = fin_rounds(ctx, n, 0, 2 * fixed_pairs, true, face_ids, st);
  } else {
    rc = fin_rounds_host(ctx, n, 0, fin_first_burst(ctx), nullptr, face_ids, st);
  }
  if (rc) return rc;
  return fin_materialize(ctx, n, face_ids, prev_scores, neighbor_list, neighbor_count,
                         used_count, fsc, st);
}

extern "C" {

// fos_jfa_faces, and with ema: the FSC EMA in the materialize launch and the
// zero-FSC count the warm-start gate needs (ctx->flags[6]).
static int jfa_faces_impl(fos_ctx *ctx, int res, int n, const int32_t *filtered, const float *rad,
                          float voxel, int32_t *face_ids, int32_t *face_counts,
                          float *prev_scores, int32_t *neighbor_list, int32_t *neighbor_count,
                          int32_t *used_count, int32_t *fsc, float *ema, float alpha,
                          hipStream_t st) {
  int *zero_cnt = ema ? ctx->flags + 6 : nullptr;
  // Emission, compaction, finalize preparation and a burst of round pairs
  // are queued together; one host read then sees both the row-table
  // overflow flag and the fixed-point flag.  An overflow regrows the tables
  // and starts over (nothing the caller keeps has been written yet: the
  // neighbour lists and hysteresis scores are written by materialize).
  while (true) {
    int rc = face_emit(ctx, res, n, filtered, st);
    if (!rc) rc = face_compact(ctx, n, face_ids, face_counts, st, true, zero_cnt);
    if (!rc) rc = fin_prepare(ctx, n, rad, voxel, nullptr, face_ids, face_counts, prev_scores, st,
                              true);
    const int first = fin_first_burst(ctx);
    if (!rc) rc = fin_rounds(ctx, n, 0, first, false, face_ids, st);
    if (rc) return rc;
    int f[64];
    rc = read_flags(ctx, f, st);
    if (rc) return rc;
    FOS_REQUIRE(!f[FOS_FLAG_FLOW],
                "pressure: a wait of the one-launch iterations timed out (results since are "
                "wrong); set pressure_flow to 0");
    if (f[0]) {
      rc = face_grow(ctx, st);
      if (rc) return rc;
      continue;
    }
    rc = fin_rounds_host(ctx, n, first, FIN_BURST, f, face_ids, st);
    if (rc) return rc;
    return fin_materialize(ctx, n, face_ids, prev_scores, neighbor_list, neighbor_count,
                           used_count, fsc, st, ema, alpha, zero_cnt);
  }
}

int fos_jfa_faces(fos_ctx *ctx, int res, int n, const int32_t *filtered, const float *rad,
                  float voxel, int32_t *face_ids, int32_t *face_counts, float *prev_scores,
                  int32_t *neighbor_list, int32_t *neighbor_count, int32_t *used_count,
                  int32_t *fsc, void *stream) {
  FOS_REQUIRE(ctx && filtered && face_ids && face_counts && res >= 1 && res <= ctx->max_res &&
                  n >= 0 && n <= ctx->max_n && voxel > 0.0f &&
                  (n == 0 || (rad && prev_scores && neighbor_list && neighbor_count &&
                              used_count && fsc)),
              "fos_jfa_faces: bad arguments (n=%d res=%d)", n, res);
  FOS_REQUIRE((double)res * res * res * 6.0 < 4294967295.0,
              "fos_jfa_faces: res=%d exceeds the 32-bit face key", res);
  if (n == 0) return 0;
  return jfa_faces_impl(ctx, res, n, filtered, rad, voxel, face_ids, face_counts, prev_scores,
                        neighbor_list, neighbor_count, used_count, fsc, nullptr, 0.0f,
                        (hipStream_t)stream);
}

int fos_jfa_faces_controller(fos_ctx *ctx, int res, int n, const int32_t *filtered,
                             float *rad_target, const float *rad, float voxel, int32_t *face_ids,
                             int32_t *face_counts, float *prev_scores, int32_t *neighbor_list,
                             int32_t *neighbor_count, int32_t *used_count, int32_t *fsc,
                             float *fsc_ema, float alpha, int frame, int warmstart_frames,
                             double fsc_zero_rate, int f_low, int f_high, float growth_pct,
                             float r_min, float r_max, float deadband, float *gate,
                             void *stream) {
  FOS_REQUIRE(ctx && filtered && face_ids && face_counts && res >= 1 && res <= ctx->max_res &&
                  n >= 1 && n <= ctx->max_n && voxel > 0.0f && rad && rad_target &&
                  prev_scores && neighbor_list && neighbor_count && used_count && fsc &&
                  fsc_ema && gate,
              "fos_jfa_faces_controller: bad arguments (n=%d res=%d)", n, res);
  FOS_REQUIRE((double)res * res * res * 6.0 < 4294967295.0,
              "fos_jfa_faces_controller: res=%d exceeds the 32-bit face key", res);
  hipStream_t st = (hipStream_t)stream;
  int rc = jfa_faces_impl(ctx, res, n, filtered, rad, voxel, face_ids, face_counts, prev_scores,
                          neighbor_list, neighbor_count, used_count, fsc, fsc_ema, alpha, st);
  if (rc) return rc;
  return fos::gate_targets(rad, rad_target, fsc, fsc_ema, n, frame, warmstart_frames,
                           fsc_zero_rate, ctx->flags + 6, f_low, f_high, growth_pct, r_min, r_max,
                           deadband, gate, st);
}

}  // extern "C"

namespace fos {

int face_step_defer(fos_ctx *ctx, const FaceStep &a, hipStream_t st) {
  FOS_REQUIRE(!ctx->spec_pending, "face_step_defer: the previous deferred face step is open");
  int *zero_cnt = ctx->flags + 6;
  int rc = face_emit(ctx, a.res, a.n, a.filtered, st);
  if (!rc) rc = face_compact(ctx, a.n, a.face_ids, a.face_counts, st, true, zero_cnt);
  if (!rc)
    rc = fin_prepare(ctx, a.n, a.rad, a.voxel, nullptr, a.face_ids, a.face_counts, a.prev_scores,
                     st, true);
  const int first = fin_first_burst(ctx);
  if (!rc) rc = fin_rounds(ctx, a.n, 0, first, false, a.face_ids, st);
  if (rc) return rc;
  if (!ctx->spec_event) {
    hipEvent_t ev;
    FOS_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    ctx->spec_event = (void *)ev;
  }
  FOS_HIP(hipMemcpyAsync(ctx->host_i, ctx->flags, 64 * sizeof(int), hipMemcpyDeviceToHost, st));
  FOS_HIP(hipEventRecord((hipEvent_t)ctx->spec_event, st));
  // held while the tables overflowed or the burst's last round changed a row
  const int *ovf = ctx->flags, *more = fin_chg(ctx) + (first - 1) % FIN_RING;
  rc = fin_materialize(ctx, a.n, a.face_ids, a.prev_scores, a.neighbor_list, a.neighbor_count,
                       a.used_count, a.fsc, st, a.fsc_ema, a.alpha, zero_cnt, ovf, more);
  if (!rc)
    rc = gate_targets(a.rad, a.rad_target, a.fsc, a.fsc_ema, a.n, a.frame, a.warmstart_frames,
                      a.fsc_zero_rate, zero_cnt, a.f_low, a.f_high, a.growth_pct, a.r_min,
                      a.r_max, a.deadband, a.gate, st, ovf, more);
  if (rc) return rc;
  ctx->spec_pending = 1;
  ctx->spec_first = first;
  return 0;
}

int face_step_finish(fos_ctx *ctx, const FaceStep &a, hipStream_t st) {
  if (!ctx->spec_pending) return 0;
  ctx->spec_pending = 0;
  FOS_HIP(hipEventSynchronize((hipEvent_t)ctx->spec_event));
  int f[64];
  const volatile int *h = (const volatile int *)ctx->host_i;
  for (int k = 0; k < 64; ++k) f[k] = h[k];
  FOS_REQUIRE(!f[FOS_FLAG_FLOW],
              "pressure: a wait of the one-launch iterations timed out (results since are "
              "wrong); set pressure_flow to 0");
  const int first = ctx->spec_first;
  int rc = 0;
  if (f[0]) {
    // row-table overflow: regrow and run the whole step (materialize and the
    // gate did nothing, so its inputs are as they were)
    rc = face_grow(ctx, st);
    if (!rc)
      rc = jfa_faces_impl(ctx, a.res, a.n, a.filtered, a.rad, a.voxel, a.face_ids, a.face_counts,
                          a.prev_scores, a.neighbor_list, a.neighbor_count, a.used_count, a.fsc,
                          a.fsc_ema, a.alpha, st);
  } else if (f[16 + (first - 1) % FIN_RING]) {
    // short of the fixed point: the rounds continue from the burst
    rc = fin_rounds_host(ctx, a.n, first, FIN_BURST, f, a.face_ids, st);
    if (!rc)
      rc = fin_materialize(ctx, a.n, a.face_ids, a.prev_scores, a.neighbor_list,
                           a.neighbor_count, a.used_count, a.fsc, st, a.fsc_ema, a.alpha,
                           ctx->flags + 6);
  } else {
    ctx->fin_burst = fin_needed(f, first - 1);
    return 0;
  }
  if (rc) return rc;
  return gate_targets(a.rad, a.rad_target, a.fsc, a.fsc_ema, a.n, a.frame, a.warmstart_frames,
                      a.fsc_zero_rate, ctx->flags + 6, a.f_low, a.f_high, a.growth_pct, a.r_min,
                      a.r_max, a.deadband, a.gate, st);
}

}  // namespace fos

extern "C" {

int fos_fsc_ema(int n, float alpha, const int32_t *fsc, float *fsc_ema,
                void *stream) {
  FOS_REQUIRE(n >= 0 && (n == 0 || (fsc && fsc_ema)), "fos_fsc_ema: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_fsc_ema, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream, n, alpha, fsc, fsc_ema);
  FOS_CHECK_LAUNCH("k_fsc_ema");
  return 0;
}

int fos_jfa_stats(fos_ctx *ctx, int n, const float *rad, float voxel,
                  const int32_t *fsc, const float *fsc_ema,
                  const int32_t *used_count, const int32_t *face_ids,
                  const int32_t *face_counts, int64_t *out_i, double *out_f,
                  void *stream) {
  FOS_REQUIRE(ctx && n >= 1 && n <= ctx->max_n && rad && fsc && fsc_ema &&
                  used_count && face_ids && face_counts && out_i && out_f,
              "fos_jfa_stats: bad arguments (n=%d)", n);
  hipStream_t st = (hipStream_t)stream;
  unsigned long long *acc = (unsigned long long *)ctx->iscratch;  // 5 x u64
  int *mm = ctx->iscratch + 10;
  double *dacc = (double *)ctx->fscratch;  // 4 doubles
  int rc = fos_fill_i32(ctx->iscratch, 10, 0, stream);
  if (rc) return rc;
  rc = fos_fill_i32(mm, 1, 0x7fffffff, stream);
  if (rc) return rc;
  rc = fos_fill_i32(mm + 1, 1, (int32_t)0x80000000, stream);
  if (rc) return rc;
  rc = fos_fill_i32((int32_t *)dacc, 8, 0, stream);
  if (rc) return rc;
  hipLaunchKernelGGL(k_stats_a, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, st, n, rad,
                     voxel, fsc, fsc_ema, used_count, face_ids, face_counts, acc, mm,
                     dacc);
  FOS_CHECK_LAUNCH("k_stats_a");
  hipLaunchKernelGGL(k_stats_b, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, st, n, rad,
                     voxel, used_count, face_ids, face_counts, dacc);
  FOS_CHECK_LAUNCH("k_stats_b");
  unsigned long long h_acc[5];
  int h_mm[2];
  double h_d[4];
  FOS_HIP(hipMemcpyAsync(h_acc, acc, sizeof(h_acc), hipMemcpyDeviceToHost, st));
  FOS_HIP(hipMemcpyAsync(h_mm, mm, sizeof(h_mm), hipMemcpyDeviceToHost, st));
  FOS_HIP(hipMemcpyAsync(h_d, dacc, sizeof(h_d), hipMemcpyDeviceToHost, st));
  FOS_HIP(hipStreamSynchronize(st));
  out_i[0] = (int64_t)h_acc[0];
  out_i[1] = (int64_t)h_acc[1];
  out_i[2] = h_mm[0];
  out_i[3] = h_mm[1];
  out_i[4] = (int64_t)h_acc[4];
  out_f[0] = (double)out_i[4] / (double)n;
  out_f[1] = h_d[0] / (double)n;
  out_f[2] = h_d[2] > 0 ? h_d[1] / h_d[2] : 0.0;
  out_f[3] = h_d[2] > 0 ? sqrt(h_d[3] / h_d[2]) : 0.0;
  return 0;
}

int fos_validate(fos_ctx *ctx, int n, const int32_t *neighbor_list,
                 const int32_t *neighbor_count, int64_t *out, void *stream) {
  FOS_REQUIRE(ctx && n >= 0 && n <= ctx->max_n && out &&
                  (n == 0 || (neighbor_list && neighbor_count)),
              "fos_validate: bad arguments (n=%d)", n);
  out[0] = out[1] = out[2] = 0;
  if (n == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  unsigned long long *acc = (unsigned long long *)ctx->iscratch;
  int rc = fos_fill_i32(ctx->iscratch, 6, 0, stream);
  if (rc) return rc;
  const int64_t m = (int64_t)n * FOS_MAX_NEIGHBORS;
  hipLaunchKernelGGL(k_validate, dim3(blocks_for(m)), dim3(FOS_BLOCK), 0, st, n,
                     neighbor_list, neighbor_count, acc);
  FOS_CHECK_LAUNCH("k_validate");
  unsigned long long h[3];
  FOS_HIP(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, st));
  FOS_HIP(hipStreamSynchronize(st));
  out[0] = (int64_t)h[0];
  out[1] = (int64_t)h[1];
  out[2] = (int64_t)h[2];
  return 0;
}

int fos_dirty_mark(const float *pos, const float *rad, const float *pos_prev,
                   const float *rad_prev, int n, int res, float L, float voxel,
                   int tile, int halo, float pos_thr, float rad_thr,
                   uint8_t *tile_dirty, void *stream) {
  FOS_REQUIRE(n >= 0 && res >= 1 && tile >= 1 && halo >= 0 && tile_dirty &&
                  (n == 0 || (pos && rad && pos_prev && rad_prev)),
              "fos_dirty_mark: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_dirty_mark, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream, pos, rad, pos_prev, rad_prev, n, res, L,
                     (float)(0.5 * (double)L), voxel, tile, halo, pos_thr, rad_thr,
                     tile_dirty);
  FOS_CHECK_LAUNCH("k_dirty_mark");
  return 0;
}

int fos_dirty_count(fos_ctx *ctx, const uint8_t *tile_dirty, int tiles,
                    int64_t *count, void *stream) {
  FOS_REQUIRE(ctx && tile_dirty && tiles >= 1 && count, "fos_dirty_count: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const int64_t m = (int64_t)tiles * tiles * tiles;
  unsigned long long *acc = (unsigned long long *)ctx->iscratch;
  int rc = fos_fill_i32(ctx->iscratch, 2, 0, stream);
  if (rc) return rc;
  hipLaunchKernelGGL(k_dirty_count, dim3(blocks_for(m)), dim3(FOS_BLOCK), 0, st,
                     tile_dirty, m, acc);
  FOS_CHECK_LAUNCH("k_dirty_count");
  unsigned long long h = 0;
  FOS_HIP(hipMemcpyAsync(&h, acc, sizeof(h), hipMemcpyDeviceToHost, st));
  FOS_HIP(hipStreamSynchronize(st));
  *count = (int64_t)h;
  return 0;
}

int fos_tile_cache_update(const float *pos, const float *rad, int n,
                          float *pos_prev, float *rad_prev, void *stream) {
  FOS_REQUIRE(n >= 0 && (n == 0 || (pos && rad && pos_prev && rad_prev)),
              "fos_tile_cache_update: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_tile_cache, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream, pos, rad, n, pos_prev, rad_prev);
  FOS_CHECK_LAUNCH("k_tile_cache");
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches of foams (no reference counterpart): every pass, the majority
// filter and the face emission run as one launch over the tiles of all foams.
// A workgroup finds its foam by binary search over the foams' first tiles
// (wave-uniform scalar loads) and then runs the single-foam tile body on that
// foam's label block.  Particle ids are batch-global, so the seed table, the
// face row tables, finalize and the FSC EMA run over the batch as one foam.
// ---------------------------------------------------------------------------

// Largest b with start[b] <= t (start ascending, start[0] = 0).
__device__ __forceinline__ int batch_find(const int *__restrict__ start, int B, int t) {
  int lo = 0, hi = B - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Step of pass p for grid edge res: res/2, halved each pass, at least 1.
__device__ __forceinline__ int batch_step(int res, int p) {
  const int s = p + 1 < 31 ? (res >> (p + 1)) : 0;
  return s < 1 ? 1 : s;
}

template <bool UNIT_L>
__global__ void k_seed_b(const float *__restrict__ pos, const float4 *__restrict__ seeds,
                         const int32_t *__restrict__ pinst, const fos_inst *__restrict__ inst,
                         int n, float L, float half_l, int32_t *__restrict__ grid) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const fos_inst I = inst[pinst[i]];
  const int res = I.res;
  int c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float nrm = (pos[3 * i + d] + half_l) / L;  // jfa.py:268-270
    c[d] = wrap_fast((int)floorf(nrm * (float)res), res);
  }
  int32_t *g = grid + I.v0;
  const int64_t v = ((int64_t)c[0] * res + c[1]) * res + c[2];
  const float cx = vcenter(c[0], I.voxel, half_l), cy = vcenter(c[1], I.voxel, half_l),
              cz = vcenter(c[2], I.voxel, half_l);
  const float pi = power_dist<UNIT_L>(cx, cy, cz, seeds[i], L);
  int cur = __hip_atomic_load(&g[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (true) {
    if (cur != -1) {
      float pc = power_dist<UNIT_L>(cx, cy, cz, seeds[cur], L);
      bool better = (pi < pc) || (pi == pc && i < cur);
      if (!better) return;
    }
    int old = atomicCAS(&g[v], cur, i);
    if (old == cur) return;
    cur = old;
  }
}

// Pass `pass` of every foam that still runs it (scheduled and not exited
// early); each foam's counters are its own stripe block.
// LT unsigned short when every foam has at most 65536 particles: list
// entries hold ids minus the foam's first id.
template <bool UNIT_L, typename LT>
__global__ void __launch_bounds__(256)
    k_jfa_pass_list_b(const float4 *__restrict__ seeds, const fos_inst *__restrict__ inst,
                      const int *__restrict__ tstart, int B, int total, int per_xcd, float L,
                      float half_l, int pass, const int32_t *__restrict__ in,
                      int32_t *__restrict__ out, int *__restrict__ stripes, int early,
                      double th, int min_passes) {
  __shared__ LT s_list[JFA_LIST][256];
  const int T = xcd_tile(blockIdx.x, per_xcd);
  if (T >= total) return;
  const int b = batch_find(tstart, B, T);
  const int res = inst[b].res;
  if (pass >= inst[b].passes) return;
  int *st = stripes + (size_t)b * FOS_BATCH_PASSES * FOS_STRIPE_PASS_INTS;
  if (!block_pass_active(st, pass, early, th, min_passes, (double)res * res * res)) return;
  const long long v0 = inst[b].v0;
  const float voxel = inst[b].voxel;
  const int nyb = (res + 3) / 4, nzb = (res + 63) / 64, ntiles = res * nyb * nzb;
  const int tile = T - tstart[b];
  const int step = batch_step(res, pass);
  const int prev = pass > 0 ? batch_step(res, pass - 1) : 0;
  const int base = inst[b].p0;
  unsigned long long bal;
  if (inst[b].tile) {
    const int tpp = (res * res + 255) / 256;
    if (2 * step == res)
      bal = jfa_flat_tile<UNIT_L, JFA_HALF, LT>(seeds, res, L, half_l, voxel, step, in + v0,
                                                out + v0, tile, tpp, s_list, base);
    else if (step == prev)
      bal = jfa_flat_tile<UNIT_L, JFA_SETTLED, LT>(seeds, res, L, half_l, voxel, step, in + v0,
                                                   out + v0, tile, tpp, s_list, base);
    else
      bal = jfa_flat_tile<UNIT_L, JFA_PLAIN, LT>(seeds, res, L, half_l, voxel, step, in + v0,
                                                 out + v0, tile, tpp, s_list, base);
  } else if (2 * step == res) {
    bal = jfa_list_tile<UNIT_L, JFA_HALF, LT>(seeds, res, L, half_l, voxel, step, in + v0,
                                              out + v0, tile, nyb, nzb, ntiles, s_list, base);
  } else if (step == prev) {
    bal = jfa_list_tile<UNIT_L, JFA_SETTLED, LT>(seeds, res, L, half_l, voxel, step, in + v0,
                                                 out + v0, tile, nyb, nzb, ntiles, s_list, base);
  } else {
    bal = jfa_list_tile<UNIT_L, JFA_PLAIN, LT>(seeds, res, L, half_l, voxel, step, in + v0,
                                               out + v0, tile, nyb, nzb, ntiles, s_list, base);
  }
  block_count(bal, st, pass);
}

// Per foam (one wave each): passes executed, exit flag and per-pass change
// counts into info[b][2 + FOS_BATCH_PASSES] (the single-foam info layout).
__global__ void k_jfa_finish_info_b(const fos_inst *__restrict__ inst, int early, double th,
                                    int min_passes, const int *__restrict__ stripes,
                                    int32_t *__restrict__ info) {
  const int b = blockIdx.x;
  const fos_inst I = inst[b];
  const double V = (double)I.res * I.res * I.res;
  const int *st = stripes + (size_t)b * FOS_BATCH_PASSES * FOS_STRIPE_PASS_INTS;
  int executed = 0;
  for (int p = 0; p < I.passes; ++p) {
    if (!pass_active(st, p, early, th, min_passes, V)) break;
    executed = p + 1;
  }
  const int fired = (early && executed > 0 && !pass_active(st, executed, early, th, min_passes, V))
                        ? 1 : 0;
  int32_t *o = info + (size_t)b * (2 + FOS_BATCH_PASSES);
  for (int p = 0; p < FOS_BATCH_PASSES; ++p) {
    const int c = p < executed ? stripes_total(st, p) : 0;
    if (threadIdx.x == 0) o[2 + p] = c;
  }
  if (threadIdx.x == 0) {
    o[0] = executed;
    o[1] = fired;
  }
}

// Foams that executed an odd number of passes hold their labels in temp:
// copy them into grid (blockIdx.y = foam).
__global__ void k_jfa_copy_odd_b(const fos_inst *__restrict__ inst,
                                 const int32_t *__restrict__ info,
                                 const int32_t *__restrict__ temp, int32_t *__restrict__ grid) {
  const int b = blockIdx.y;
  if (!(info[(size_t)b * (2 + FOS_BATCH_PASSES)] & 1)) return;
  const fos_inst I = inst[b];
  const int64_t V = (int64_t)I.res * I.res * I.res;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < V;
       i += (int64_t)gridDim.x * blockDim.x)
    grid[I.v0 + i] = temp[I.v0 + i];
}

__global__ void __launch_bounds__(256)
    k_majority_b(const fos_inst *__restrict__ inst, const int *__restrict__ tstart, int B,
                 int total, int per_xcd, const int32_t *__restrict__ g, int32_t *__restrict__ f) {
  const int T = xcd_tile(blockIdx.x, per_xcd);
  if (T >= total) return;
  const int b = batch_find(tstart, B, T);
  const int res = inst[b].res;
  const long long v0 = inst[b].v0;
  const int nyb = (res + 3) / 4, nzb = (res + 63) / 64;
  const int tile = T - tstart[b];
  const int zb = tile % nzb, t2 = tile / nzb;
  const int yb = t2 % nyb, x = t2 / nyb;
  const int z = zb * 64 + (int)(threadIdx.x & 63), y = yb * 4 + (int)(threadIdx.x >> 6);
  if (z >= res || y >= res) return;
  majority_vox(res, g + v0, f + v0, x, y, z);
}

__global__ void __launch_bounds__(256)
    k_face_emit_b(const fos_inst *__restrict__ inst, const int *__restrict__ estart, int B,
                  int total, int per_xcd, int n, const int32_t *__restrict__ f, int *tb,
                  unsigned *tk, int *tc, int cap, int *overflow) {
  __shared__ unsigned long long s_pair[EMIT_HS];
  __shared__ unsigned s_key[EMIT_HS];
  __shared__ int s_cnt[EMIT_HS];
  const int T = xcd_tile(blockIdx.x, per_xcd);
  if (T >= total) return;
  const int b = batch_find(estart, B, T);
  const int res = inst[b].res;
  const int nyb = (res + 256 / EMIT_B_ZW - 1) / (256 / EMIT_B_ZW),
            nzb = (res + EMIT_B_ZW - 1) / EMIT_B_ZW;
  const int ntiles = ((res + EMIT_B_TX - 1) / EMIT_B_TX) * nyb * nzb;
  face_emit_tile<EMIT_B_TX, EMIT_B_ZW>(T - estart[b], res, n, f + inst[b].v0, tb, tk, tc, cap,
                                       overflow, nyb, nzb, ntiles, s_pair, s_key, s_cnt);
}

extern "C" {

int fos_batch_jfa_init(fos_batch *bt, fos_ctx *ctx, const float *pos, const float *rad,
                       float L, float beta, int32_t *grid, int32_t *temp, void *stream) {
  FOS_REQUIRE(bt && ctx && grid && pos && rad && L > 0.0f && bt->n_total <= ctx->max_n,
              "fos_batch_jfa_init: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  int rc = fos_fill_i32(grid, bt->v_total, -1, stream);
  if (rc) return rc;
  if (temp) {
    rc = fos_fill_i32(temp, bt->v_total, -1, stream);
    if (rc) return rc;
  }
  const int n = bt->n_total;
  if (n == 0) return 0;
  rc = fos_jfa_seed_table(ctx, pos, rad, n, beta, stream);
  if (rc) return rc;
  const float half_l = (float)(0.5 * (double)L);
  if (unit_l(L))
    hipLaunchKernelGGL(k_seed_b<true>, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, st, pos,
                       ctx->seeds, bt->pinst, bt->inst, n, L, half_l, grid);
  else
    hipLaunchKernelGGL(k_seed_b<false>, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, st, pos,
                       ctx->seeds, bt->pinst, bt->inst, n, L, half_l, grid);
  FOS_CHECK_LAUNCH("k_seed_b");
  return 0;
}

int fos_batch_jfa_passes(fos_batch *bt, fos_ctx *ctx, float L, int early_exit,
                         double th_pass_delta, int min_passes, int32_t *grid, int32_t *temp,
                         int32_t *info, void *stream) {
  FOS_REQUIRE(bt && ctx && grid && temp && info && grid != temp && L > 0.0f &&
                  bt->max_passes <= FOS_BATCH_PASSES - 1,
              "fos_batch_jfa_passes: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const int B = bt->B;
  int rc = fos_fill_i32(bt->stripes, (int64_t)B * FOS_BATCH_PASSES * FOS_STRIPE_PASS_INTS, 0,
                        stream);
  if (rc) return rc;
  const int total = bt->tstart_h[B];
  const int per = (total + FOS_XCDS - 1) / FOS_XCDS;
  const float half_l = (float)(0.5 * (double)L);
  bool small = true;  // every foam's ids fit 16 bits above its first id
  for (int b = 0; b < B; ++b) small = small && bt->inst_h[b].n <= 65536;
  for (int p = 0; p < bt->max_passes; ++p) {
    const int32_t *in = (p & 1) ? temp : grid;
    int32_t *out = (p & 1) ? grid : temp;
#define FOS_PASS_B(U, T)                                                                   \
  hipLaunchKernelGGL((k_jfa_pass_list_b<U, T>), dim3(per * FOS_XCDS), dim3(256), 0, st,       \
                     ctx->seeds, bt->inst, bt->tstart, B, total, per, L, half_l, p, in, out,   \
                     bt->stripes, early_exit, th_pass_delta, min_passes)
    if (unit_l(L)) {
      if (small) FOS_PASS_B(true, unsigned short); else FOS_PASS_B(true, int);
    } else {
      if (small) FOS_PASS_B(false, unsigned short); else FOS_PASS_B(false, int);
    }
#undef FOS_PASS_B
    FOS_CHECK_LAUNCH("k_jfa_pass_list_b");
  }
  hipLaunchKernelGGL(k_jfa_finish_info_b, dim3(B), dim3(64), 0, st, bt->inst, early_exit,
                     th_pass_delta, min_passes, (const int *)bt->stripes, info);
  FOS_CHECK_LAUNCH("k_jfa_finish_info_b");
  hipLaunchKernelGGL(k_jfa_copy_odd_b, dim3(256, B), dim3(FOS_BLOCK), 0, st, bt->inst,
                     (const int32_t *)info, temp, grid);
  FOS_CHECK_LAUNCH("k_jfa_copy_odd_b");
  return 0;
}

int fos_batch_majority(fos_batch *bt, const int32_t *grid, int32_t *filtered, void *stream) {
  FOS_REQUIRE(bt && grid && filtered && grid != filtered, "fos_batch_majority: bad arguments");
  const int total = bt->mstart_h[bt->B];
  const int per = (total + FOS_XCDS - 1) / FOS_XCDS;
  hipLaunchKernelGGL(k_majority_b, dim3(per * FOS_XCDS), dim3(256), 0, (hipStream_t)stream,
                     bt->inst, bt->mstart, bt->B, total, per, grid, filtered);
  FOS_CHECK_LAUNCH("k_majority_b");
  return 0;
}

int fos_batch_face_accumulate(fos_batch *bt, fos_ctx *ctx, const int32_t *filtered,
                              int32_t *face_ids, int32_t *face_counts, int check,
                              void *stream) {
  FOS_REQUIRE(bt && ctx && filtered && face_ids && face_counts && bt->n_total <= ctx->max_n,
              "fos_batch_face_accumulate: bad arguments");
  FOS_REQUIRE((double)bt->max_res * bt->max_res * bt->max_res * 6.0 < 4294967295.0,
              "fos_batch_face_accumulate: res=%d exceeds the 32-bit face key", bt->max_res);
  const int n = bt->n_total;
  if (n == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  const int total = bt->estart_h[bt->B];
  const int per = (total + FOS_XCDS - 1) / FOS_XCDS;
  while (true) {
    hipLaunchKernelGGL(k_face_emit_b, dim3(per * FOS_XCDS), dim3(256), 0, st, bt->inst,
                       bt->estart, bt->B, total, per, n, filtered, ctx->tab_b, ctx->tab_k,
                       ctx->tab_c, ctx->tab_cap, ctx->flags);
    FOS_CHECK_LAUNCH("k_face_emit_b");
    if (!check) break;  // the caller reads the overflow flag (fos_batch_flags)
    FOS_HIP(hipMemcpyAsync(ctx->host_i, ctx->flags, sizeof(int), hipMemcpyDeviceToHost, st));
    FOS_HIP(hipStreamSynchronize(st));
    if (!*(volatile int *)ctx->host_i) break;
    const int cap = ctx->tab_cap * 4;
    FOS_REQUIRE(cap <= 4096, "fos_batch_face_accumulate: row table overflow at %d",
                ctx->tab_cap);
    int rc = alloc_tables(ctx, cap);
    if (rc) return rc;
    rc = fos_fill_i32(ctx->flags, 1, 0, stream);
    if (rc) return rc;
  }
  const int cap = ctx->tab_cap;
  const size_t shmem = (size_t)4 * cap * sizeof(unsigned);
  hipLaunchKernelGGL(k_face_compact, dim3((unsigned)((n + 3) / 4)), dim3(256), shmem, st, n,
                     ctx->tab_b, ctx->tab_k, ctx->tab_c, cap, face_ids, face_counts,
                     (int *)nullptr, (int *)nullptr, (unsigned *)nullptr, (int *)nullptr, 0,
                     (int *)nullptr, (int *)nullptr, (int *)nullptr);
  FOS_CHECK_LAUNCH("k_face_compact");
  return 0;
}

int fos_batch_finalize_faces(fos_batch *bt, fos_ctx *ctx, const float *rad,
                             const int32_t *face_ids, const int32_t *face_counts,
                             float *prev_scores, int32_t *neighbor_list, int32_t *neighbor_count,
                             int32_t *used_count, int32_t *fsc, int fixed_pairs, void *stream) {
  FOS_REQUIRE(bt && ctx && bt->n_total <= ctx->max_n &&
                  (bt->n_total == 0 || (rad && face_ids && face_counts && prev_scores &&
                                        neighbor_list && neighbor_count && used_count && fsc)),
              "fos_batch_finalize_faces: bad arguments");
  FOS_REQUIRE(fixed_pairs >= 0, "fos_batch_finalize_faces: fixed_pairs %d", fixed_pairs);
  return finalize_impl(ctx, bt->n_total, rad, 1.0f, bt->pvox, face_ids, face_counts,
                       prev_scores, neighbor_list, neighbor_count, used_count, fsc, fixed_pairs,
                       stream);
}

int fos_batch_flags(fos_ctx *ctx, int32_t *out, void *stream) {
  FOS_REQUIRE(ctx && out, "fos_batch_flags: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  FOS_HIP(hipMemcpyAsync(ctx->host_i, ctx->flags, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
  FOS_HIP(hipStreamSynchronize(st));
  const int *h = (const int *)ctx->host_i;
  out[0] = h[0];
  out[1] = h[3];
  return 0;
}

}  // extern "C"
