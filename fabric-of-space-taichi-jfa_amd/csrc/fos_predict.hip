// fos_predict.hip -- learned face predictor: scores hash-grid candidate pairs
// for a shared power-diagram face on frames where the JFA is decimated.
//
// No reference counterpart: the reference always runs the JFA (jfa.run_jfa,
// jfa.py:1056-1197).  Candidate generation follows the shape of
// topology.build_topo_candidates (topology.py:68-133: 27 hash cells, a
// capped list per particle), here for every j != i.
//
// Pipeline per frame: candidates -> 13 features per (i, j) (12 geometric,
// one hysteresis) -> MLP 13 -> 64 -> 64 -> 1 (BF16 weights and activations
// on MFMA, FP32 accumulation, ReLU) -> logit > 0 is a face (sigmoid 0.5) ->
// FSC and face rows.  The training step (the same GEMMs on MFMA, forward and
// backward, with Adam in FP32) fits the MLP to labels taken from the plain
// JFA's neighbour lists.
#include "fos_common.h"

using namespace fos;

#define PRED_FEAT 16   // feature slots per pair: PRED_NFEAT used, the rest zero (MFMA K = 16)
#define PRED_NFEAT 13
#define MLP_H 64
// flat FP32 parameter layout: W1[16][64] b1[64] W2[64][64] b2[64] w3[64] b3
#define MLP_OFF_W1 0
#define MLP_OFF_B1 (16 * MLP_H)
#define MLP_OFF_W2 (MLP_OFF_B1 + MLP_H)
#define MLP_OFF_B2 (MLP_OFF_W2 + MLP_H * MLP_H)
#define MLP_OFF_W3 (MLP_OFF_B2 + MLP_H)
#define MLP_OFF_B3 (MLP_OFF_W3 + MLP_H)
#define MLP_NPARAM (MLP_OFF_B3 + 1)

__device__ __forceinline__ float dot3f(float x, float y, float z) {
  return (x * x + y * y) + z * z;
}

// ---------------------------------------------------------------------------
// Candidates: every j != i with |d(i, j)| < mult * (ri + rj), scanning the 27
// cells around i's hash cell (dx, dy, dz order, ascending index inside a
// cell), at most `cap` kept in that order.  A row lists the partners j > i
// first (nup[i] of them, in scan order), then the partners j < i.  Needs a
// hash cell of at least mult * 2 * r_max.
// ---------------------------------------------------------------------------
#define PRED_CAP_MAX 128   // row capacity limit (LDS staging of the j < i block)
#define PRED_NB_MAX 1024   // particles of a 27-cell neighbourhood staged in LDS

// Row state of one particle while a wave scans its neighbourhood 64 hits at
// a time: partners j > i go straight to the front of the row, partners
// j < i to the wave's LDS buffers (copied behind them by cand_finish).
// Next to each kept partner k the row of `rel` (when given) holds
// (pk - pi wrapped to the minimum image, |pk - pi|^2 - rk^2).
struct CandRow {
  int cu, cl, tr;
};

__device__ __forceinline__ void cand_emit(CandRow &st, bool hit, int j, float4 r, int i,
                                          int lane, int cap, int32_t *__restrict__ row,
                                          float4 *__restrict__ relrow,
                                          int32_t *__restrict__ low, float4 *__restrict__ lowr) {
  const unsigned long long below = (1ull << lane) - 1ull;
  const unsigned long long mh = __ballot(hit);
  const bool keep = hit && st.cu + st.cl + __popcll(mh & below) < cap;
  const unsigned long long mk = __ballot(keep);
  const unsigned long long mu = __ballot(keep && j > i);
  if (keep) {
    if (j > i) {
      const int o = st.cu + __popcll(mu & below);
      row[o] = j;
      if (relrow) relrow[o] = r;
    } else {
      const int o = st.cl + __popcll(mk & ~mu & below);
      low[o] = j;
      lowr[o] = r;
    }
  }
  st.tr |= (mh & ~mk) != 0ull;
  st.cu += __popcll(mu);
  st.cl += __popcll(mk & ~mu);
}

__device__ __forceinline__ void cand_finish(const CandRow &st, int i, int lane, int cap,
                                            int32_t *__restrict__ row,
                                            float4 *__restrict__ relrow,
                                            const int32_t *__restrict__ low,
                                            const float4 *__restrict__ lowr,
                                            int32_t *__restrict__ ncand,
                                            int32_t *__restrict__ nup,
                                            int32_t *__restrict__ trunc) {
  __builtin_amdgcn_wave_barrier();
  for (int t = lane; t < st.cl; t += 64) {
    row[st.cu + t] = low[t];
    if (relrow) relrow[st.cu + t] = lowr[t];
  }
  for (int t = st.cu + st.cl + lane; t < cap; t += 64) row[t] = -1;
  if (lane == 0) {
    ncand[i] = st.cu + st.cl;
    nup[i] = st.cu;
    trunc[i] = st.tr;
  }
}

// The hit test of the candidate kernels and the rel entry of a hit.
__device__ __forceinline__ bool cand_test(float qx, float qy, float qz, float rk, float px,
                                          float py, float pz, float ri, float L, float inv_l,
                                          float mult, float4 &r) {
  const float ddx = gwrap(qx - px, L, inv_l);
  const float ddy = gwrap(qy - py, L, inv_l);
  const float ddz = gwrap(qz - pz, L, inv_l);
  const float d2 = dot3f(ddx, ddy, ddz);
  const float th = mult * (ri + rk);
  r = make_float4(ddx, ddy, ddz, d2 - rk * rk);
  return d2 < th * th;
}

// One workgroup per hash cell: the 27-cell neighbourhood (cell/occupant
// order) is staged in LDS as (wrapped position, radius) and particle ids,
// then each wave scans it for the cell's particles in turn.  A
// neighbourhood of more than PRED_NB_MAX particles is left to
// k_pred_cand_wave (its particles get ncand = -1 here).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
    k_pred_cand_cell(const float *__restrict__ pos, const float *__restrict__ rad, int G,
                     float L, float inv_l, float mult, int cap,
                     const int32_t *__restrict__ cs, const int32_t *__restrict__ cc,
                     const int32_t *__restrict__ ci, int32_t *__restrict__ cand,
                     int32_t *__restrict__ ncand, int32_t *__restrict__ nup,
                     float4 *__restrict__ rel, int32_t *__restrict__ trunc) {
  __shared__ float4 s_p[PRED_NB_MAX];
  __shared__ int32_t s_j[PRED_NB_MAX];
  __shared__ int32_t s_low[4][PRED_CAP_MAX];
  __shared__ float4 s_lowr[4][PRED_CAP_MAX];
  __shared__ int32_t s_beg[27], s_off[28];
  const int c = (int)blockIdx.x;
  const int nc = cc[c];
  if (nc == 0) return;
  const int tid = (int)threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int c0 = cs[c];
  if (w == 0) {
    const int cx = c / (G * G), cy = (c / G) % G, cz = c % G;
    int b = 0, m = 0;
    if (lane < 27) {
      const int id = (wrap1(cx + lane / 9 - 1, G) * G + wrap1(cy + (lane / 3) % 3 - 1, G)) * G +
                     wrap1(cz + lane % 3 - 1, G);
      b = cs[id];
      m = cc[id];
    }
    int x = m;  // inclusive prefix over the 27 cells
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
      const int y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane < 27) {
      s_beg[lane] = b;
      s_off[lane + 1] = x;
    }
    if (lane == 0) s_off[0] = 0;
  }
  __syncthreads();
  const int T = s_off[27];
  if (T > PRED_NB_MAX) {
    for (int p = tid; p < nc; p += 256) ncand[ci[c0 + p]] = -1;
    return;
  }
  for (int t = tid; t < T; t += 256) {
    int q = 0;  // cell of slot t: s_off[q] <= t < s_off[q + 1]
#pragma unroll
    for (int h = 16; h >= 1; h >>= 1)
      if (q + h < 28 && s_off[q + h] <= t) q += h;
    const int j = ci[s_beg[q] + (t - s_off[q])];
    s_j[t] = j;
    s_p[t] = make_float4(pos[3 * j], pos[3 * j + 1], pos[3 * j + 2], rad[j]);
  }
  __syncthreads();
  for (int p = w; p < nc; p += 4) {
    const int i = ci[c0 + p];
    const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2], ri = rad[i];
    CandRow st = {0, 0, 0};
    int32_t *row = cand + (int64_t)i * cap;
    float4 *relrow = rel ? rel + (int64_t)i * cap : nullptr;
    for (int t0 = 0; t0 < T; t0 += 64) {
      const int t = t0 + lane;
      bool hit = false;
      int j = -1;
      float4 r = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (t < T) {
        j = s_j[t];
        if (j != i) {
          const float4 q = s_p[t];
          hit = cand_test(q.x, q.y, q.z, q.w, px, py, pz, ri, L, inv_l, mult, r);
        }
      }
      cand_emit(st, hit, j, r, i, lane, cap, row, relrow, s_low[w], s_lowr[w]);
    }
    cand_finish(st, i, lane, cap, row, relrow, s_low[w], s_lowr[w], ncand, nup, trunc);
    __builtin_amdgcn_wave_barrier();  // the LDS buffers are reused by the next particle
  }
}

// One wave per particle over the 27 cells in global memory; only for
// particles the cell kernel left (ncand[i] < 0).
__global__ void __launch_bounds__(256)
    k_pred_cand_wave(const float *__restrict__ pos, const float *__restrict__ rad, int n, int G,
                     float L, float half_l, float inv_l, float mult, int cap,
                     const int32_t *__restrict__ cs, const int32_t *__restrict__ cc,
                     const int32_t *__restrict__ ci, int32_t *__restrict__ cand,
                     int32_t *__restrict__ ncand, int32_t *__restrict__ nup,
                     float4 *__restrict__ rel, int32_t *__restrict__ trunc) {
  __shared__ int32_t s_low[4][PRED_CAP_MAX];
  __shared__ float4 s_lowr[4][PRED_CAP_MAX];
  const int lane = (int)(threadIdx.x & 63);
  const int w = (int)(threadIdx.x >> 6);
  const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (i >= n || ncand[i] >= 0) return;
  const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2], ri = rad[i];
  // the particle's own hash cell, as the grid rebuild computed it
  const int cx = hash_axis(gwrap(px, L, inv_l), G, half_l, inv_l);
  const int cy = hash_axis(gwrap(py, L, inv_l), G, half_l, inv_l);
  const int cz = hash_axis(gwrap(pz, L, inv_l), G, half_l, inv_l);
  CandRow st = {0, 0, 0};
  int32_t *row = cand + (int64_t)i * cap;
  float4 *relrow = rel ? rel + (int64_t)i * cap : nullptr;
  for (int q = 0; q < 27; ++q) {
    const int id = (wrap1(cx + q / 9 - 1, G) * G + wrap1(cy + (q / 3) % 3 - 1, G)) * G +
                   wrap1(cz + q % 3 - 1, G);
    const int b = cs[id], e = b + cc[id];
    for (int k0 = b; k0 < e; k0 += 64) {
      const int k = k0 + lane;
      bool hit = false;
      int j = -1;
      float4 r = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (k < e) {
        j = ci[k];
        if (j != i)
          hit = cand_test(pos[3 * j], pos[3 * j + 1], pos[3 * j + 2], rad[j], px, py, pz, ri, L,
                          inv_l, mult, r);
      }
      cand_emit(st, hit, j, r, i, lane, cap, row, relrow, s_low[w], s_lowr[w]);
    }
  }
  cand_finish(st, i, lane, cap, row, relrow, s_low[w], s_lowr[w], ncand, nup, trunc);
}

// Flat list of the scored pairs (i, cand[i][s]), s < nup[i]: pair p belongs
// to particle pair_i[p] and slot p - off[i]; off[n] is the number of pairs.
__global__ void k_pred_pair_map(const int32_t *__restrict__ nup, int n, int32_t *__restrict__ off,
                                int32_t *__restrict__ pair_i) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = off[i], m = nup[i];
  for (int s2 = 0; s2 < m; ++s2) pair_i[o + s2] = i;
  if (i == n - 1) off[n] = o + m;
}

// ---------------------------------------------------------------------------
// Features of the directed pair (i, j) = (i, cand[i][s]).  With d = pj - pi
// (minimum image), the power plane of i and j cuts the segment at
// t = (|d|^2 + ri^2 - rj^2) / (2 |d|^2); m = pi + t d has equal power
// p(m) = |m - pi|^2 - ri^2 for i and j.  A candidate k of i with
// |m - pk|^2 - rk^2 < p(m) owns m, so it may cut the face off.
// The face itself is estimated as a square of half-size clip_r * (ri + rj)
// around m in the power plane, clipped by the half-planes p_i(x) <= p_k(x)
// of the PRED_NEAR candidates k of i whose lines pass closest to m; the area
// comes from a half-plane intersection with fixed trip counts.  A voxelised
// face of area A and unit normal n crosses about A (|nx| + |ny| + |nz|) /
// voxel^2 voxel faces, so A (|nx| + |ny| + |nz|) / (1.5 (ri + rj)^2)
// estimates the JFA's face score (jfa.py:379-397 expected_area).
//   0 |d| / (ri + rj)         6 witnesses among i's candidates / 4 (<= 8)
//   1 ri / r_mean             7 min witness margin / (ri + rj)^2
//   2 rj / r_mean             8 |cand(i)| / 40
//   3 t                       9 |cand(j)| / 40
//   4 estimated score (<= 4) 10 (ri - rj) / (ri + rj)
//   5 face area / (ri + rj)^2 (<= 4)   11 p(m) / (ri + rj)^2
//  12 1 when j was among i's accepted partners j > i of the previous frame
//     (prev_up rows; the JFA's hysteresis keeps faces accepted before,
//     jfa.py:400-443), else 0
// Margins and p(m) are clamped to [-1, 1].  The candidate relation is
// symmetric, so each unordered pair is scored once, in the row of its lower
// index: the pairs are the flat list of k_pred_pair_map, one thread each,
// reading i's candidates (relative positions and |pk - pi|^2 - rk^2) from
// the rel rows the candidate kernels wrote.
// ---------------------------------------------------------------------------
#define PRED_NEAR 8    // planes kept for the face polygon (nearest to m)

// Six waves per SIMD (at most 85 VGPRs; 112 unbounded): C5 198 -> 184 us,
// the same features.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
    k_pred_feat(const float *__restrict__ rad, int n, float inv_rbar, float clip_r,
                const int32_t *__restrict__ cand, const int32_t *__restrict__ ncand,
                const float4 *__restrict__ rel, const int32_t *__restrict__ off,
                const int32_t *__restrict__ pair_i, int cap, int64_t max_pairs,
                const int32_t *__restrict__ prev_up, float *__restrict__ feat) {
  const int64_t P = min((int64_t)off[n], max_pairs);
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < P;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int i = pair_i[p];
    const int s = (int)(p - off[i]);
    const int64_t rb = (int64_t)i * cap;
    const int j = cand[rb + s];
    const int ni = ncand[i];
    const float4 *sk = rel + rb;
    const float ri = rad[i], rj = rad[j];
    const float4 ks = sk[s];
    float4 kn = sk[0];  // next candidate, loaded one iteration ahead
    const float dx = ks.x, dy = ks.y, dz = ks.z;
    const float d2 = fmaxf(dot3f(dx, dy, dz), 1e-20f);
    const float sr = ri + rj, sr2 = sr * sr;
    const float t = (d2 + ri * ri - rj * rj) / (2.0f * d2);
    const float pm = t * t * d2 - ri * ri;
    const float ri2 = ri * ri;
    // orthonormal basis (u, v) of the power plane
    const float dl = sqrtf(d2);
    const float nx = dx / dl, ny = dy / dl, nz = dz / dl;
    float ux, uy, uz;
    if (fabsf(nx) < 0.9f) {
      ux = 1.0f - nx * nx; uy = -nx * ny; uz = -nx * nz;
    } else {
      ux = -ny * nx; uy = 1.0f - ny * ny; uz = -ny * nz;
    }
    const float ul = rsqrtf(dot3f(ux, uy, uz));
    ux *= ul; uy *= ul; uz *= ul;
    const float vx = ny * uz - nz * uy, vy = nz * ux - nx * uz, vz = nx * uy - ny * ux;
    // nearest PRED_NEAR planes, ascending offset: the insertion keeps only
    // (offset, candidate) and the unit normals of the winners are recomputed
    // afterwards with the same expressions (same bits, fewer swaps)
    float pa[PRED_NEAR], pb[PRED_NEAR], pc[PRED_NEAR];
    int pq[PRED_NEAR];
#pragma unroll
    for (int h = 0; h < PRED_NEAR; ++h) {
      pa[h] = 1.0f; pb[h] = 0.0f; pc[h] = 1e30f; pq[h] = -1;
    }
    bool empty = false;
    int wit = 0;
    float mw = 1e30f;
    const float mm = t * t * d2 - pm;  // |m|^2 - p(m) = ri^2
    for (int q = 0; q < ni; ++q) {
      const float4 kq = kn;
      if (q + 1 < ni) kn = sk[q + 1];
      if (q == s) continue;
      const float kx = kq.x, ky = kq.y, kz = kq.z, pk = kq.w;
      const float kd = kx * dx + ky * dy + kz * dz;
      // |pk - m|^2 - rk^2 - p(m), with m = t d
      const float wm = pk - 2.0f * t * kd + mm;
      wit += wm < 0.0f ? 1 : 0;
      mw = fminf(mw, wm);
      // p_i(x) <= p_k(x) on the plane, x = m + a u + b v:  A a + B b <= C
      const float A = 2.0f * (kx * ux + ky * uy + kz * uz);
      const float B = 2.0f * (kx * vx + ky * vy + kz * vz);
      const float C = pk + ri2 - 2.0f * t * kd;
      const float q2 = A * A + B * B;
      if (q2 < 1e-30f) {
        empty |= C < 0.0f;
        continue;
      }
      const float inv = rsqrtf(q2);
      float ic = C * inv;
      if (!(ic < pc[PRED_NEAR - 1])) continue;
      int iq = q;
#pragma unroll
      for (int h = 0; h < PRED_NEAR; ++h) {
        const bool sw = ic < pc[h];
        const float tc = pc[h];
        const int tq = pq[h];
        pc[h] = sw ? ic : tc; pq[h] = sw ? iq : tq;
        ic = sw ? tc : ic; iq = sw ? tq : iq;
      }
    }
#pragma unroll
    for (int h = 0; h < PRED_NEAR; ++h) {
      if (pq[h] < 0) continue;
      const float4 kq = sk[pq[h]];
      const float kx = kq.x, ky = kq.y, kz = kq.z;
      const float A = 2.0f * (kx * ux + ky * uy + kz * uz);
      const float B = 2.0f * (kx * vx + ky * vy + kz * vz);
      const float inv = rsqrtf(A * A + B * B);
      pa[h] = A * inv;
      pb[h] = B * inv;
    }
    // Area of {x : a_h x <= c_h} over the kept planes and the four sides of a
    // square of half-size clip_r*(ri+rj): each line contributes the edge it
    // keeps after clipping by all the others, summed with the shoelace formula
    // (lines traversed with the interior on the left).
    const float R0 = clip_r * sr;
    float la[PRED_NEAR + 4], lb[PRED_NEAR + 4], lc[PRED_NEAR + 4];
#pragma unroll
    for (int h = 0; h < PRED_NEAR; ++h) {
      la[h] = pa[h]; lb[h] = pb[h]; lc[h] = pc[h];
    }
    la[PRED_NEAR] = 1.0f;  lb[PRED_NEAR] = 0.0f;  lc[PRED_NEAR] = R0;
    la[PRED_NEAR + 1] = -1.0f; lb[PRED_NEAR + 1] = 0.0f; lc[PRED_NEAR + 1] = R0;
    la[PRED_NEAR + 2] = 0.0f; lb[PRED_NEAR + 2] = 1.0f;  lc[PRED_NEAR + 2] = R0;
    la[PRED_NEAR + 3] = 0.0f; lb[PRED_NEAR + 3] = -1.0f; lc[PRED_NEAR + 3] = R0;
    float area = 0.0f;
#pragma unroll
    for (int h = 0; h < PRED_NEAR + 4; ++h) {
      const float x0 = la[h] * lc[h], y0 = lb[h] * lc[h];  // closest point to m
      const float tx = -lb[h], ty = la[h];
      float lo = -1e30f, hi = 1e30f;
#pragma unroll
      for (int g = 0; g < PRED_NEAR + 4; ++g) {
        if (g == h) continue;
        const float kk = la[g] * tx + lb[g] * ty;
        const float rr = lc[g] - (la[g] * x0 + lb[g] * y0);
        const float e = rr * __builtin_amdgcn_rcpf(kk);  // edge parameter, ~1 ulp
        if (kk > 1e-12f)
          hi = fminf(hi, e);
        else if (kk < -1e-12f)
          lo = fmaxf(lo, e);
        else if (rr < 0.0f || (rr == 0.0f && g < h))
          hi = -1e30f;  // parallel and tighter (or an equal line met first)
      }
      if (lc[h] < 1e29f && lo < hi) {
        const float x1 = x0 + lo * tx, y1 = y0 + lo * ty;
        const float x2 = x0 + hi * tx, y2 = y0 + hi * ty;
        area += x1 * y2 - x2 * y1;
      }
    }
    area = empty ? 0.0f : fmaxf(0.0f, 0.5f * area);
    const float na = area / sr2;
    float f[PRED_FEAT];
    f[0] = dl / sr;
    f[1] = ri * inv_rbar;
    f[2] = rj * inv_rbar;
    f[3] = t;
    f[4] = fminf(4.0f, na * (fabsf(nx) + fabsf(ny) + fabsf(nz)) / 1.5f);
    f[5] = fminf(4.0f, na);
    f[6] = (float)min(wit, 8) * 0.25f;
    f[7] = fminf(1.0f, fmaxf(-1.0f, mw / sr2));
    f[8] = (float)ni * (1.0f / 40.0f);
    f[9] = (float)ncand[j] * (1.0f / 40.0f);
    f[10] = (ri - rj) / sr;
    f[11] = fminf(1.0f, fmaxf(-1.0f, pm / sr2));
    float was = 0.0f;
    if (prev_up) {
      const int32_t *pr = prev_up + (int64_t)i * FOS_MAX_NEIGHBORS;
#pragma unroll 8
      for (int q = 0; q < FOS_MAX_NEIGHBORS; ++q) was = pr[q] == j ? 1.0f : was;
    }
    f[12] = was;
    f[13] = f[14] = f[15] = 0.0f;
    float4 *o = (float4 *)(feat + p * PRED_FEAT);
#pragma unroll
    for (int q = 0; q < PRED_FEAT / 4; ++q)
      o[q] = make_float4(f[4 * q], f[4 * q + 1], f[4 * q + 2], f[4 * q + 3]);
  }
}

// ---------------------------------------------------------------------------
// MLP inference on MFMA (v_mfma_f32_16x16x16_bf16).  Each layer is computed
// transposed, Y^T = W^T X^T: the weights are the A operand (pre-packed on the
// host in lane order), the activations the B operand.  The 16x16 f32 result
// of a layer leaves lane l holding rows 4*(l/16)..+3 of column l%16, which
// is the B-operand layout of the next layer's 16-row K block, so hidden
// activations stay in registers.  A wave scores 16 pairs per step.  The
// pair count is npairs, or min(npairs, *npairs_dev) when that is given.
// ---------------------------------------------------------------------------
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// fp32 -> bf16 bits, round to nearest even (gfx950: v_cvt_pk_bf16_f32)
__device__ __forceinline__ short f2bf(float x) { return __builtin_bit_cast(short, (__bf16)x); }

__global__ void __launch_bounds__(256)
    k_mlp_infer(const float *__restrict__ feat, int64_t npairs,
                const bf16x4 *__restrict__ w1p, const bf16x4 *__restrict__ w2p,
                const float *__restrict__ b1, const float *__restrict__ b2,
                const float *__restrict__ w3, float b3, float *__restrict__ logit,
                const int32_t *__restrict__ npairs_dev) {
  const int lane = (int)(threadIdx.x & 63);
  const int col = lane & 15, grp = lane >> 4;
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
  const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  bf16x4 a1[4], a2[4][4];
  float bb1[4][4], bb2[4][4], ww3[4][4];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    a1[mb] = w1p[mb * 64 + lane];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) a2[mb][kb] = w2p[(mb * 4 + kb) * 64 + lane];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int h = 16 * mb + 4 * grp + v;
      bb1[mb][v] = b1[h];
      bb2[mb][v] = b2[h];
      ww3[mb][v] = w3[h];
    }
  }
  const int64_t P = npairs_dev ? min(npairs, (int64_t)*npairs_dev) : npairs;
  for (int64_t p0 = wave * 16; p0 < P; p0 += waves * 16) {
    const int64_t pc = p0 + col;
    const float4 x = pc < P ? *(const float4 *)(feat + pc * PRED_FEAT + 4 * grp)
                            : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    bf16x4 xb;
    xb[0] = f2bf(x.x);
    xb[1] = f2bf(x.y);
    xb[2] = f2bf(x.z);
    xb[3] = f2bf(x.w);
    bf16x4 h1[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1[mb], xb, acc, 0, 0, 0);
#pragma unroll
      for (int v = 0; v < 4; ++v) h1[mb][v] = f2bf(fmaxf(acc[v] + bb1[mb][v], 0.0f));
    }
    float part = 0.0f;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb)
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2[mb][kb], h1[kb], acc, 0, 0, 0);
#pragma unroll
      for (int v = 0; v < 4; ++v) part += fmaxf(acc[v] + bb2[mb][v], 0.0f) * ww3[mb][v];
    }
    part += __shfl_xor(part, 16, 64);
    part += __shfl_xor(part, 32, 64);
    if (grp == 0 && pc < P) logit[pc] = part + b3;
  }
}

// FSC from the scores: the pair p = (i, cand[i][s]), s = p - off[i], is kept
// when logit[p] > 0 (sigmoid 0.5) and then counts for both particles (fsc
// must be zero on entry).  Face rows list the kept partners j > i in
// candidate order, at most 32, -1 padded (the row shape accumulate_face_area
// produces for the pressure step).  One wave per i.
__global__ void __launch_bounds__(256)
    k_pred_fsc(const float *__restrict__ logit, const int32_t *__restrict__ cand,
               const int32_t *__restrict__ off, int n, int cap, int32_t *__restrict__ fsc,
               int32_t *__restrict__ face_ids, int32_t *__restrict__ face_counts) {
  const int lane = (int)(threadIdx.x & 63);
  const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (i >= n) return;
  const int o = off[i], nu = off[i + 1] - o;
  const unsigned long long below = (1ull << lane) - 1ull;
  int c = 0;
  for (int s0 = 0; s0 < nu; s0 += 64) {
    const int s = s0 + lane;
    bool keep = false;
    int j = -1;
    if (s < nu) {
      keep = logit[o + s] > 0.0f;
      j = cand[(int64_t)i * cap + s];
    }
    const unsigned long long mk = __ballot(keep);
    if (keep) {
      atomicAdd(&fsc[j], 1);
      const int r = c + __popcll(mk & below);
      if (face_ids && r < FOS_MAX_NEIGHBORS) {
        face_ids[(int64_t)i * FOS_MAX_NEIGHBORS + r] = j;
        if (face_counts) face_counts[(int64_t)i * FOS_MAX_NEIGHBORS + r] = 1;
      }
    }
    c += __popcll(mk);
  }
  if (lane == 0 && c) atomicAdd(&fsc[i], c);
  if (face_ids)
    for (int r = min(c, FOS_MAX_NEIGHBORS) + lane; r < FOS_MAX_NEIGHBORS; r += 64) {
      face_ids[(int64_t)i * FOS_MAX_NEIGHBORS + r] = -1;
      if (face_counts) face_counts[(int64_t)i * FOS_MAX_NEIGHBORS + r] = 0;
    }
}

// Row i of prev_up: the partners j > i of i's neighbour list, in list order,
// -1 padded (the accepted pairs of a JFA frame, each in its lower row).
__global__ void k_pred_prev_rows(const int32_t *__restrict__ nl, const int32_t *__restrict__ nc,
                                 int n, int32_t *__restrict__ prev_up) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int m = min(nc[i], FOS_MAX_NEIGHBORS);
  int r = 0;
  for (int q = 0; q < m; ++q) {
    const int j = nl[(int64_t)i * FOS_MAX_NEIGHBORS + q];
    if (j > i) prev_up[(int64_t)i * FOS_MAX_NEIGHBORS + r++] = j;
  }
  for (; r < FOS_MAX_NEIGHBORS; ++r) prev_up[(int64_t)i * FOS_MAX_NEIGHBORS + r] = -1;
}

__global__ void k_fill_i32(int32_t *__restrict__ a, int n, int32_t v) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) a[q] = v;
}

// Labels for training, one per pair p = (i, cand[i][s]): 1 when cand[i][s]
// is in i's JFA neighbour list, else 0.
__global__ void k_pred_labels(const int32_t *__restrict__ cand, const int32_t *__restrict__ off,
                              const int32_t *__restrict__ pair_i, int n, int cap,
                              int64_t max_pairs, const int32_t *__restrict__ nl,
                              const int32_t *__restrict__ nc, float *__restrict__ y) {
  const int64_t P = min((int64_t)off[n], max_pairs);
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < P;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int i = pair_i[p];
    const int j = cand[(int64_t)i * cap + (p - off[i])];
    const int m = min(nc[i], FOS_MAX_NEIGHBORS);
    float v = 0.0f;
    for (int q = 0; q < m; ++q)
      if (nl[(int64_t)i * FOS_MAX_NEIGHBORS + q] == j) v = 1.0f;
    y[p] = v;
  }
}

// ---------------------------------------------------------------------------
// Training on MFMA (v_mfma_f32_16x16x16_bf16, FP32 accumulation).  One step
// over a batch of B samples (indices into the feature/label arrays) is two
// launches:
//   k_mlp_train  one wave per workgroup; wave w takes the 16-sample tiles
//                w, w + W, w + 2W, ...  Per tile, in registers:
//                  forward   Z1^T = W1^T X^T, H1 = relu(Z1 + b1),
//                            Z2^T = W2^T H1^T, H2 = relu(Z2 + b2),
//                            z3 = H2 w3 + b3 (the inference kernel's layout);
//                  backward  dz3 = w (sigmoid(z3) - t) / B,
//                            dZ2 = [Z2 + b2 > 0] dz3 w3^T,
//                            dZ1^T = [Z1 + b1 > 0] (W2 dZ2^T);
//                  gradients gW2^T += dZ2^T H1, gW1^T += dZ1^T X,
//                            with the batch as the MFMA K dimension: H1^T,
//                            dZ2^T, dZ1^T and X^T go through LDS to put the
//                            sample index in the K slots;
//                            b1, b2, w3, b3 and the loss in FP32 VALU.
//                The wave writes its partial sums (and loss) to part[w].
//   k_mlp_adam   sums the W partials per parameter in wave order, then Adam.
// X, W1, W2, H1, dZ2 and dZ1 enter the MFMAs as BF16 (round to nearest
// even, as the inference kernel rounds X, W and H1); Z1, Z2, H2, dz3, dZ2,
// dZ1 and every sum are FP32.  Feature slots PRED_NFEAT..15 are read as 0,
// so rows PRED_NFEAT..15 of W1 get a zero gradient.  The result does not
// depend on the launch beyond W, and W depends only on B, so a step is
// reproducible bit for bit.
// Loss: class-weighted binary cross entropy on the logit.
// ---------------------------------------------------------------------------
#define MLP_TRAIN_WAVES 64
#define MLP_PART (MLP_NPARAM + 1)   // per-wave partial: the gradient, then the loss sum
static_assert(MLP_TRAIN_WAVES * MLP_PART == FOS_MLP_TRAIN_SCRATCH, "fos.h scratch size");

__global__ void __launch_bounds__(64)
    k_mlp_train(const float *__restrict__ feat, const float *__restrict__ y,
                const int32_t *__restrict__ idx, int B, const float *__restrict__ prm,
                float pos_w, float *__restrict__ part) {
  // transposes: rows are a hidden unit (or feature), columns the 16 samples
  __shared__ __align__(16) short sH1[MLP_H][16], sD2[MLP_H][16], sD1[MLP_H][16], sX[16][16];
  const int lane = (int)threadIdx.x;
  const int col = lane & 15, grp = lane >> 4;
  const int w = (int)blockIdx.x, W = (int)gridDim.x;
  // weights as A operands: W1^T and W2^T blocks for the forward (lane holds
  // W[16kb + 4grp + v][16mb + col]), W2 blocks for the backward (lane holds
  // W2[16ma + col][16kb + 4grp + v])
  bf16x4 a1[4], a2[4][4], a2b[4][4];
  float bb1[4][4], bb2[4][4], ww3[4][4];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int k = 4 * grp + v, h = 16 * mb + k;
      a1[mb][v] = f2bf(prm[MLP_OFF_W1 + k * MLP_H + 16 * mb + col]);
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) {
        a2[mb][kb][v] = f2bf(prm[MLP_OFF_W2 + (16 * kb + k) * MLP_H + 16 * mb + col]);
        a2b[mb][kb][v] = f2bf(prm[MLP_OFF_W2 + (16 * mb + col) * MLP_H + 16 * kb + k]);
      }
      bb1[mb][v] = prm[MLP_OFF_B1 + h];
      bb2[mb][v] = prm[MLP_OFF_B2 + h];
      ww3[mb][v] = prm[MLP_OFF_W3 + h];
    }
  }
  const float b3 = prm[MLP_OFF_B3];
  const float inv_b = 1.0f / (float)B;
  f32x4 gw2[4][4], gw1[4];
  float gb1[4][4], gb2[4][4], gw3[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    gw1[i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      gw2[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      gb1[i][j] = gb2[i][j] = gw3[i][j] = 0.0f;
    }
  }
  float gb3 = 0.0f, loss = 0.0f;
  const int tiles = (B + 15) >> 4;
  for (int t = w; t < tiles; t += W) {
    const int s = 16 * t + col;
    const bool live = s < B;
    const int64_t p = live ? (int64_t)idx[s] : 0;
    float4 x = live ? *(const float4 *)(feat + p * PRED_FEAT + 4 * grp)
                    : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    static_assert(PRED_NFEAT == 13, "feature slots 13..15 are the zero ones");
    if (grp == 3) x.y = x.z = x.w = 0.0f;
    bf16x4 xb;
    xb[0] = f2bf(x.x);
    xb[1] = f2bf(x.y);
    xb[2] = f2bf(x.z);
    xb[3] = f2bf(x.w);
    // forward
    bf16x4 h1b[4];
    unsigned m1 = 0u;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1[mb], xb, acc, 0, 0, 0);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float z = acc[v] + bb1[mb][v];
        if (z > 0.0f) m1 |= 1u << (4 * mb + v);
        h1b[mb][v] = f2bf(fmaxf(z, 0.0f));
      }
    }
    float z2[4][4];
    float part3 = 0.0f;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb)
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2[mb][kb], h1b[kb], acc, 0, 0, 0);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        z2[mb][v] = acc[v] + bb2[mb][v];
        part3 += fmaxf(z2[mb][v], 0.0f) * ww3[mb][v];
      }
    }
    // the four lanes of a sample add the same four partials (same bits)
    part3 += __shfl_xor(part3, 16, 64);
    part3 += __shfl_xor(part3, 32, 64);
    const float z3 = part3 + b3;
    const float tg = live ? y[p] : 0.0f;
    const float wt = tg > 0.5f ? pos_w : 1.0f;
    const float dz3 = live ? wt * (1.0f / (1.0f + expf(-z3)) - tg) * inv_b : 0.0f;
    if (grp == 0 && live) {
      // log(1 + e^-|z|) + max(z, 0) - t z
      loss += wt * (log1pf(expf(-fabsf(z3))) + fmaxf(z3, 0.0f) - tg * z3);
      gb3 += dz3;
    }
    // backward
    bf16x4 d2b[4], d1b[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float d = z2[mb][v] > 0.0f ? dz3 * ww3[mb][v] : 0.0f;
        gb2[mb][v] += d;
        gw3[mb][v] += dz3 * fmaxf(z2[mb][v], 0.0f);
        d2b[mb][v] = f2bf(d);
      }
    }
#pragma unroll
    for (int ma = 0; ma < 4; ++ma) {
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb)
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2b[ma][kb], d2b[kb], acc, 0, 0, 0);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float d = (m1 >> (4 * ma + v)) & 1u ? acc[v] : 0.0f;
        gb1[ma][v] += d;
        d1b[ma][v] = f2bf(d);
      }
    }
    // transposes through LDS: lane (col, grp) holds rows 16mb + 4grp + v of
    // sample col and reads back row 16mb + col, samples 4grp..4grp+3
    __syncthreads();
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = 16 * mb + 4 * grp + v;
        sH1[r][col] = h1b[mb][v];
        sD2[r][col] = d2b[mb][v];
        sD1[r][col] = d1b[mb][v];
      }
#pragma unroll
    for (int v = 0; v < 4; ++v) sX[4 * grp + v][col] = xb[v];
    __syncthreads();
    bf16x4 tH1[4], tD2[4], tD1[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      tH1[mb] = *(const bf16x4 *)&sH1[16 * mb + col][4 * grp];
      tD2[mb] = *(const bf16x4 *)&sD2[16 * mb + col][4 * grp];
      tD1[mb] = *(const bf16x4 *)&sD1[16 * mb + col][4 * grp];
    }
    const bf16x4 tX = *(const bf16x4 *)&sX[col][4 * grp];
    // gW2^T[16mh + 4grp + v][16na + col] and gW1^T[16ma + 4grp + v][col]
#pragma unroll
    for (int mh = 0; mh < 4; ++mh)
#pragma unroll
      for (int na = 0; na < 4; ++na)
        gw2[mh][na] =
            __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(tD2[mh], tH1[na], gw2[mh][na], 0, 0, 0);
#pragma unroll
    for (int ma = 0; ma < 4; ++ma)
      gw1[ma] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(tD1[ma], tX, gw1[ma], 0, 0, 0);
  }
  float *out = part + (int64_t)w * MLP_PART;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int h = 16 * mb + 4 * grp + v;
      out[MLP_OFF_W1 + col * MLP_H + h] = gw1[mb][v];
#pragma unroll
      for (int na = 0; na < 4; ++na) out[MLP_OFF_W2 + (16 * na + col) * MLP_H + h] = gw2[mb][na][v];
      // the 16 sample lanes of a row group, in a fixed butterfly
      float s1 = gb1[mb][v], s2 = gb2[mb][v], s3 = gw3[mb][v];
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) {
        s1 += __shfl_xor(s1, o, 64);
        s2 += __shfl_xor(s2, o, 64);
        s3 += __shfl_xor(s3, o, 64);
      }
      if (col == 0) {
        out[MLP_OFF_B1 + h] = s1;
        out[MLP_OFF_B2 + h] = s2;
        out[MLP_OFF_W3 + h] = s3;
      }
    }
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    gb3 += __shfl_xor(gb3, o, 64);
    loss += __shfl_xor(loss, o, 64);
  }
  if (lane == 0) {
    out[MLP_OFF_B3] = gb3;
    out[MLP_NPARAM] = loss;
  }
}

// Sum of the W partials per parameter (wave order) into grad, then Adam;
// slot MLP_NPARAM adds the batch mean loss to *loss.
__global__ void k_mlp_adam(const float *__restrict__ part, int W, float *__restrict__ grad,
                           float *__restrict__ prm, float *__restrict__ m, float *__restrict__ v,
                           float lr, float b1, float b2, float eps, float c1, float c2,
                           float inv_b, float *__restrict__ loss) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > MLP_NPARAM) return;
  float g = 0.0f;
  for (int k = 0; k < W; ++k) g += part[(int64_t)k * MLP_PART + q];
  if (q == MLP_NPARAM) {
    loss[0] += g * inv_b;
    return;
  }
  grad[q] = g;
  const float mq = b1 * m[q] + (1.0f - b1) * g;
  const float vq = b2 * v[q] + (1.0f - b2) * g * g;
  m[q] = mq;
  v[q] = vq;
  prm[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

extern "C" {

int fos_pred_candidates(const float *pos, const float *rad, int n, int G, float L, float mult,
                        int cap, const int32_t *cell_start, const int32_t *cell_count,
                        const int32_t *cell_indices, int32_t *cand, int32_t *ncand, int32_t *nup,
                        float *rel, int32_t *truncated, void *stream) {
  FOS_REQUIRE(n >= 0 && G >= 3 && L > 0.0f && cap >= 1 && mult > 0.0f &&
                  (n == 0 || (pos && rad && cell_start && cell_count && cell_indices && cand &&
                              ncand && nup && truncated)),
              "fos_pred_candidates: bad arguments (n=%d G=%d cap=%d)", n, G, cap);
  FOS_REQUIRE(cap <= PRED_CAP_MAX, "fos_pred_candidates: cap %d exceeds %d", cap, PRED_CAP_MAX);
  FOS_REQUIRE((int64_t)G * G * G <= 0x7fffffff, "fos_pred_candidates: G=%d too large", G);
  if (n == 0) return 0;
  const float inv_l = (float)(1.0 / (double)L);
  hipLaunchKernelGGL(k_pred_cand_cell, dim3((unsigned)(G * G * G)), dim3(256), 0,
                     (hipStream_t)stream, pos, rad, G, L, inv_l, mult, cap, cell_start,
                     cell_count, cell_indices, cand, ncand, nup, (float4 *)rel, truncated);
  FOS_CHECK_LAUNCH("k_pred_cand_cell");
  hipLaunchKernelGGL(k_pred_cand_wave, dim3(blocks_for((int64_t)n * 64)), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream,
                     pos, rad, n, G, L, (float)(0.5 * (double)L), inv_l, mult,
                     cap, cell_start, cell_count, cell_indices, cand, ncand, nup, (float4 *)rel,
                     truncated);
  FOS_CHECK_LAUNCH("k_pred_cand_wave");
  return 0;
}

int fos_pred_pairs(fos_ctx *ctx, const int32_t *nup, int n, int32_t *pair_off, int32_t *pair_i,
                   void *stream) {
  FOS_REQUIRE(ctx && n >= 0 && (n == 0 || (nup && pair_off && pair_i)),
              "fos_pred_pairs: bad arguments (n=%d)", n);
  if (n == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  if (exclusive_scan(ctx, nup, pair_off, n, st)) return -1;
  hipLaunchKernelGGL(k_pred_pair_map, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, st, nup, n,
                     pair_off, pair_i);
  FOS_CHECK_LAUNCH("k_pred_pair_map");
  return 0;
}

static unsigned pair_blocks(int64_t max_pairs) {
  const int64_t b = (max_pairs + 255) / 256;
  return (unsigned)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

int fos_pred_features(const float *rad, int n, float r_mean, float clip_r, const int32_t *cand,
                      const int32_t *ncand, const float *rel, const int32_t *pair_off,
                      const int32_t *pair_i, int cap, int64_t max_pairs, const int32_t *prev_up,
                      float *feat, void *stream) {
  FOS_REQUIRE(n >= 0 && cap >= 1 && r_mean > 0.0f && clip_r > 0.0f && max_pairs >= 0 &&
                  (n == 0 || (rad && cand && ncand && rel && pair_off && pair_i && feat)),
              "fos_pred_features: bad arguments (n=%d cap=%d)", n, cap);
  if (n == 0 || max_pairs == 0) return 0;
  hipLaunchKernelGGL(k_pred_feat, dim3(pair_blocks(max_pairs)), dim3(256), 0,
                     (hipStream_t)stream, rad, n, 1.0f / r_mean, clip_r, cand, ncand,
                     (const float4 *)rel, pair_off, pair_i, cap, max_pairs, prev_up, feat);
  FOS_CHECK_LAUNCH("k_pred_feat");
  return 0;
}

int fos_pred_prev_rows(const int32_t *neighbor_list, const int32_t *neighbor_count, int n,
                       int32_t *prev_up, void *stream) {
  FOS_REQUIRE(n >= 0 && (n == 0 || (neighbor_list && neighbor_count && prev_up)),
              "fos_pred_prev_rows: bad arguments (n=%d)", n);
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_pred_prev_rows, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream, neighbor_list, neighbor_count, n, prev_up);
  FOS_CHECK_LAUNCH("k_pred_prev_rows");
  return 0;
}

int fos_pred_mlp(const float *feat, int64_t npairs, const void *w1_packed, const void *w2_packed,
                 const float *b1, const float *b2, const float *w3, float b3, float *logit,
                 const int32_t *npairs_dev, void *stream) {
  FOS_REQUIRE(npairs >= 0 &&
                  (npairs == 0 || (feat && w1_packed && w2_packed && b1 && b2 && w3 && logit)),
              "fos_pred_mlp: bad arguments (npairs=%lld)", (long long)npairs);
  if (npairs == 0) return 0;
  const int64_t groups = (npairs + 15) / 16;          // one wave step per 16 pairs
  int64_t blocks = (groups + 3) / 4;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     feat, npairs, (const bf16x4 *)w1_packed, (const bf16x4 *)w2_packed, b1, b2,
                     w3, b3, logit, npairs_dev);
  FOS_CHECK_LAUNCH("k_mlp_infer");
  return 0;
}

int fos_pred_fsc(const float *logit, const int32_t *cand, const int32_t *pair_off, int n, int cap,
                 int32_t *fsc, int32_t *face_ids, int32_t *face_counts, void *stream) {
  FOS_REQUIRE(n >= 0 && cap >= 1 && (n == 0 || (logit && cand && pair_off && fsc)),
              "fos_pred_fsc: bad arguments (n=%d cap=%d)", n, cap);
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, (hipStream_t)stream,
                     fsc, n, 0);
  FOS_CHECK_LAUNCH("k_fill_i32");
  hipLaunchKernelGGL(k_pred_fsc, dim3(blocks_for((int64_t)n * 64)), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream, logit, cand, pair_off, n, cap, fsc, face_ids,
                     face_counts);
  FOS_CHECK_LAUNCH("k_pred_fsc");
  return 0;
}

int fos_pred_labels(const int32_t *cand, const int32_t *pair_off, const int32_t *pair_i, int n,
                    int cap, int64_t max_pairs, const int32_t *neighbor_list,
                    const int32_t *neighbor_count, float *labels, void *stream) {
  FOS_REQUIRE(n >= 0 && cap >= 1 && max_pairs >= 0 &&
                  (n == 0 || (cand && pair_off && pair_i && neighbor_list && neighbor_count &&
                              labels)),
              "fos_pred_labels: bad arguments (n=%d cap=%d)", n, cap);
  if (n == 0 || max_pairs == 0) return 0;
  hipLaunchKernelGGL(k_pred_labels, dim3(pair_blocks(max_pairs)), dim3(256), 0,
                     (hipStream_t)stream, cand, pair_off, pair_i, n, cap, max_pairs,
                     neighbor_list, neighbor_count, labels);
  FOS_CHECK_LAUNCH("k_pred_labels");
  return 0;
}

int fos_mlp_param_count(int *count) {
  FOS_REQUIRE(count, "fos_mlp_param_count: count is NULL");
  *count = MLP_NPARAM;
  return 0;
}

int fos_mlp_train_step(float *params, float *grad, float *adam_m, float *adam_v, int step,
                       float lr, const float *feat, const float *labels, const int32_t *batch_idx,
                       int B, float pos_weight, float *scratch, float *loss, void *stream) {
  FOS_REQUIRE(params && grad && adam_m && adam_v && feat && labels && batch_idx && scratch &&
                  loss && B >= 1 && step >= 1 && lr > 0.0f,
              "fos_mlp_train_step: bad arguments (B=%d step=%d)", B, step);
  hipStream_t st = (hipStream_t)stream;
  const int tiles = (B + 15) / 16;
  const int W = tiles < MLP_TRAIN_WAVES ? tiles : MLP_TRAIN_WAVES;
  hipLaunchKernelGGL(k_mlp_train, dim3(W), dim3(64), 0, st, feat, labels, batch_idx, B,
                     (const float *)params, pos_weight, scratch);
  FOS_CHECK_LAUNCH("k_mlp_train");
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_mlp_adam, dim3(blocks_for(MLP_PART)), dim3(FOS_BLOCK), 0, st,
                     (const float *)scratch, W, grad, params, adam_m, adam_v, lr, b1, b2, 1e-8f,
                     c1, c2, 1.0f / (float)B, loss);
  FOS_CHECK_LAUNCH("k_mlp_adam");
  return 0;
}

}  // extern "C"
