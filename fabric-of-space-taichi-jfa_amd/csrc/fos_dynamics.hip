// fos_dynamics.hip -- PBD overlap projection, max overlap, FSC controller,
// pressure equilibration, Brownian/OU noise, XSPH, damping and the topology
// candidate builder (reference: dynamics.py, topology.py).
#include <string.h>

#include "fos_common.h"

using namespace fos;

// Walks the 27 hash cells around particle i in the reference order
// (dx, dy, dz in -1..1, occupants in cell order) and calls f(j).
template <typename F>
__device__ __forceinline__ void for_stencil(int cx, int cy, int cz, int G,
                                            const int32_t *__restrict__ cs,
                                            const int32_t *__restrict__ cc,
                                            const int32_t *__restrict__ ci, F f) {
  for (int dx = -1; dx <= 1; ++dx) {
    const int nx = wrap_fast(cx + dx, G);
    for (int dy = -1; dy <= 1; ++dy) {
      const int ny = wrap_fast(cy + dy, G);
      for (int dz = -1; dz <= 1; ++dz) {
        const int nz = wrap_fast(cz + dz, G);
        const int id = nx * G * G + ny * G + nz;
        const int st = cs[id], ct = cc[id];
        for (int k = st; k < st + ct; ++k) f(ci[k]);
      }
    }
  }
}

__device__ __forceinline__ float dot3(float x, float y, float z) {
  float d2 = x * x + y * y;
  return d2 + z * z;
}

// PBD with 32 lanes per particle (8 particles per 256-thread block): lane c <
// 27 walks stencil cell c = (dx+1)*9 + (dy+1)*3 + (dz+1) and parks its
// overlap corrections in LDS in occupant order; lane 0 then adds them in the
// reference order (cells dx, dy, dz, occupants ascending), so the sum is the
// serial one bit for bit.  A cell with more than PBD_SLOTS corrections sends
// its particle through the serial loop instead.
#define PBD_SLOTS 6
__device__ __forceinline__ void pbd_pair(int j, int i, float px, float py, float pz,
                                         float ri, float maxd, float gap, float L,
                                         float inv_l, const float *__restrict__ pin,
                                         const float *__restrict__ rad, bool &hit,
                                         float &vx, float &vy, float &vz) {
  hit = false;
  if (j == i) return;
  const float dx = gwrap(pin[3 * j] - px, L, inv_l);
  const float dy = gwrap(pin[3 * j + 1] - py, L, inv_l);
  const float dz = gwrap(pin[3 * j + 2] - pz, L, inv_l);
  const float dist = sqrtf(dot3(dx, dy, dz) + 1e-12f);
  const float ts = (ri + rad[j]) * gap;
  if (dist < ts) {
    const float overlap = ts - dist;
    const float ux = dx / dist, uy = dy / dist, uz = dz / dist;
    const float mag = fminf(overlap * 0.5f, maxd);
    vx = fmaxf(-maxd, fminf(maxd, -ux * mag));
    vy = fmaxf(-maxd, fminf(maxd, -uy * mag));
    vz = fmaxf(-maxd, fminf(maxd, -uz * mag));
    hit = true;
  }
}

// BATCH: particles of a batch of foams; each uses its foam's hash grid
// (edge G, cells from c0) and sits the pass out (copies its position) when
// its foam runs fewer than pass + 1 passes this frame.
template <bool BATCH>
__global__ void __launch_bounds__(256)
    k_pbd_wave(const float *__restrict__ pin, float *__restrict__ pout,
               const float *__restrict__ rad, int n, int G, float L, float half_l,
               float inv_l, float gap, float mdf, float dmul,
               const int32_t *__restrict__ cs, const int32_t *__restrict__ cc,
               const int32_t *__restrict__ ci, const int32_t *__restrict__ pinst,
               const fos_inst *__restrict__ inst, const int32_t *__restrict__ npass, int pass,
               int32_t *__restrict__ cell_of) {
  __shared__ float s_v[8][27][PBD_SLOTS][3];
  __shared__ int s_cnt[8][27];
  const int grp = (int)(threadIdx.x >> 5), l = (int)(threadIdx.x & 31);
  const int i = blockIdx.x * 8 + grp;
  bool live = i < n;
  if (BATCH && live) {
    const int b = pinst[i];
    if (pass >= npass[b]) {
      if (l < 3) pout[3 * i + l] = pin[3 * i + l];
      live = false;
    } else {
      G = inst[b].G;
      cs += inst[b].c0;
      cc += inst[b].c0;
    }
  }
  float px = 0.f, py = 0.f, pz = 0.f, ri = 0.f, maxd = 0.f;
  int cx = 0, cy = 0, cz = 0;
  if (live) {
    px = pin[3 * i];
    py = pin[3 * i + 1];
    pz = pin[3 * i + 2];
    ri = rad[i];
    maxd = mdf * ri;
    cx = centre_axis(px, G, L, half_l, inv_l);
    cy = centre_axis(py, G, L, half_l, inv_l);
    cz = centre_axis(pz, G, L, half_l, inv_l);
  }
  if (live && l < 27) {
    const int id = wrap_fast(cx + l / 9 - 1, G) * G * G + wrap_fast(cy + (l / 3) % 3 - 1, G) * G +
                   wrap_fast(cz + l % 3 - 1, G);
    const int st = cs[id], ct = cc[id];
    int m = 0;
    for (int k = st; k < st + ct; ++k) {
      bool hit;
      float vx, vy, vz;
      pbd_pair(ci[k], i, px, py, pz, ri, maxd, gap, L, inv_l, pin, rad, hit, vx, vy, vz);
      if (hit) {
        if (m < PBD_SLOTS) {
          s_v[grp][l][m][0] = vx;
          s_v[grp][l][m][1] = vy;
          s_v[grp][l][m][2] = vz;
        }
        m += 1;
      }
    }
    s_cnt[grp][l] = m;
  }
  __syncthreads();
  if (!live || l != 0) return;
  bool spill = false;
  for (int c = 0; c < 27; ++c) spill |= s_cnt[grp][c] > PBD_SLOTS;
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
  if (!spill) {
    for (int c = 0; c < 27; ++c)
      for (int q = 0; q < s_cnt[grp][c]; ++q) {
        ax = ax + s_v[grp][c][q][0];
        ay = ay + s_v[grp][c][q][1];
        az = az + s_v[grp][c][q][2];
      }
  } else {
    for_stencil(cx, cy, cz, G, cs, cc, ci, [&](int j) {
      bool hit;
      float vx, vy, vz;
      pbd_pair(j, i, px, py, pz, ri, maxd, gap, L, inv_l, pin, rad, hit, vx, vy, vz);
      if (hit) {
        ax = ax + vx;
        ay = ay + vy;
        az = az + vz;
      }
    });
  }
  const float maxt = maxd * dmul;
  const float cm = sqrtf(dot3(ax, ay, az));
  if (cm > maxt) {
    const float s = maxt / cm;
    ax = ax * s;
    ay = ay * s;
    az = az * s;
  }
  float w0 = gwrap(px + ax, L, inv_l), w1 = gwrap(py + ay, L, inv_l),
        w2 = gwrap(pz + az, L, inv_l);
  if (!BATCH && cell_of) {
    // the cell step of the rebuild that follows (grid.py:285-309): wrap
    // again and hash
    w0 = gwrap(w0, L, inv_l);
    w1 = gwrap(w1, L, inv_l);
    w2 = gwrap(w2, L, inv_l);
    cell_of[i] = hash_axis(w0, G, half_l, inv_l) * G * G + hash_axis(w1, G, half_l, inv_l) * G +
                 hash_axis(w2, G, half_l, inv_l);
  }
  pout[3 * i] = w0;
  pout[3 * i + 1] = w1;
  pout[3 * i + 2] = w2;
}

// Max overlap with 32 lanes per particle: the maximum is order-free, so the
// lanes' partial maxima combine with a half-wave reduction.
// BATCH: each particle uses its foam's hash grid and its foam's maximum
// (gmax[b]).
template <bool BATCH>
__global__ void __launch_bounds__(256)
    k_max_overlap_wave(const float *__restrict__ pos, const float *__restrict__ rad, int n,
                       int G, float L, float half_l, float inv_l,
                       const int32_t *__restrict__ cs, const int32_t *__restrict__ cc,
                       const int32_t *__restrict__ ci, float *__restrict__ lmd,
                       float *gmax, const int32_t *__restrict__ pinst,
                       const fos_inst *__restrict__ inst) {
  const int grp = (int)(threadIdx.x >> 5), l = (int)(threadIdx.x & 31);
  const int i = blockIdx.x * 8 + grp;
  float lm = 0.0f;
  if (BATCH && i < n) {
    const int b = pinst[i];
    G = inst[b].G;
    cs += inst[b].c0;
    cc += inst[b].c0;
    gmax += b;
  }
  if (i < n && l < 27) {
    const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
    const float ri = rad[i];
    const int cx = centre_axis(px, G, L, half_l, inv_l);
    const int cy = centre_axis(py, G, L, half_l, inv_l);
    const int cz = centre_axis(pz, G, L, half_l, inv_l);
    const int id = wrap_fast(cx + l / 9 - 1, G) * G * G + wrap_fast(cy + (l / 3) % 3 - 1, G) * G +
                   wrap_fast(cz + l % 3 - 1, G);
    for (int k = cs[id]; k < cs[id] + cc[id]; ++k) {
      const int j = ci[k];
      if (!(i < j)) continue;
      const float dx = gwrap(pos[3 * j] - px, L, inv_l);
      const float dy = gwrap(pos[3 * j + 1] - py, L, inv_l);
      const float dz = gwrap(pos[3 * j + 2] - pz, L, inv_l);
      const float dist = sqrtf(dot3(dx, dy, dz) + 1e-8f);
      const float td = ri + rad[j];
      if (dist < td) lm = fmaxf(lm, td - dist);
    }
  }
  for (int o = 16; o > 0; o >>= 1) lm = fmaxf(lm, __shfl_xor(lm, o, 32));
  if (i < n && l == 0) {
    lmd[i] = lm;
    if (BATCH && lm > 0.0f) atomicMax((int *)gmax, __float_as_int(lm));
  }
  if (!BATCH) {
    // one atomic per block on the single maximum, not one per particle
    // (atomics on one address serialise); lm >= 0, so the int image of the
    // float orders as the float does
    __shared__ float s_m[8];
    if (l == 0) s_m[grp] = (i < n) ? lm : 0.0f;
    __syncthreads();
    if (threadIdx.x == 0) {
      float m = s_m[0];
#pragma unroll
      for (int g = 1; g < 8; ++g) m = fmaxf(m, s_m[g]);
      if (m > 0.0f) atomicMax((int *)gmax, __float_as_int(m));
    }
  }
}

// dynamics.py:1226-1240
__device__ __forceinline__ float smoothstep(float e0, float e1, float x) {
  const float t = fminf(1.0f, fmaxf(0.0f, (x - e0) / (e1 - e0)));
  return t * t * (3.0f - 2.0f * t);
}

// dynamics.py:1281-1320
// gate (may be NULL): run only when *gate > 0.5 (the warm-start gate decided
// on the device, fos_fsc_gate).
__global__ void k_fsc_targets(const float *__restrict__ rad, float *__restrict__ rt,
                              const int32_t *__restrict__ fsc,
                              const float *__restrict__ ema, int n, float fl,
                              float fh, float gp, float r_min, float r_max,
                              float db, const float *__restrict__ gate) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (gate && !(*gate > 0.5f))) return;
  const float r0 = rad[i];
  const float f = ema[i] > 0.0f ? ema[i] : (float)fsc[i];
  const float gap = 1.0f, idle = 0.003f;
  float out;
  if (f < fl - gap) {
    out = fminf(r0 * (1.0f + gp * 1.0f), r_max);
  } else if (f < fl) {
    const float g = 1.0f - smoothstep(fl - gap - db, fl, f);
    out = fminf(r0 * (1.0f + gp * g), r_max);
  } else if (f > fh + gap) {
    out = fmaxf(r0 * (1.0f - gp * 1.0f), r_min);
  } else if (f > fh) {
    const float g = smoothstep(fh, fh + gap + db, f);
    out = fmaxf(r0 * (1.0f - gp * g), r_min);
  } else {
    const float center = 0.5f * (fl + fh);
    if (f < center)
      out = fminf(r0 * (1.0f + gp * idle), r_max);
    else
      out = fmaxf(r0 * (1.0f - gp * idle), r_min);
  }
  rt[i] = out;
}

// dynamics.py:1345-1360: counts, sum of dr (double), max |dr| (float bits)
__global__ void k_controller_stats(const float *__restrict__ rad,
                                   const float *__restrict__ rt,
                                   const float *__restrict__ ema, int n, float fl,
                                   float fh, unsigned long long *cnt, double *sum,
                                   int *maxbits) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float f = ema[i];
  const float dr = rt[i] - rad[i];
  if (f < fl)
    atomicAdd(&cnt[0], 1ull);
  else if (f > fh)
    atomicAdd(&cnt[1], 1ull);
  else
    atomicAdd(&cnt[2], 1ull);
  atomicAdd(sum, (double)dr);
  const float a = fabsf(dr);
  if (a > 0.0f) atomicMax(maxbits, __float_as_int(a));
}

// run.py:1184-1195 warm-start gate on the device: *gate = 1 when the frame is
// past the warm start and at most zero_rate of the particles have FSC 0.
// Count and decision in one launch: every block adds its zeros to cnt[0];
// the last block to finish (cnt[1] counts finished blocks) decides the gate
// and resets both counters for the next call.
__global__ void __launch_bounds__(FOS_BLOCK)
    k_count_gate(const int32_t *__restrict__ fsc, int n, int frame, int warmstart,
                 double zero_rate, int *__restrict__ cnt, float *__restrict__ gate) {
  __shared__ int s_last;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = __popcll(__ballot(i < n && fsc[i] == 0));
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(&cnt[0], c);
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) s_last = atomicAdd(&cnt[1], 1) == (int)gridDim.x - 1;
  __syncthreads();
  if (!s_last || threadIdx.x != 0) return;
  __threadfence();
  const int zeros = atomicAdd(&cnt[0], 0);
  const double f0 = (double)zeros / (double)(n > 1 ? n : 1);
  *gate = (frame < warmstart || f0 > zero_rate) ? 0.0f : 1.0f;
  atomicExch(&cnt[0], 0);
  atomicExch(&cnt[1], 0);
}

// run.py:1184-1195 gate from a complete zero-FSC count, then (gate open)
// dynamics.py:1281-1320 set_fsc_targets: one launch where fos_fsc_gate and
// fos_fsc_targets_gated take two.  Every thread derives the same gate in
// double; block 0 thread 0 publishes it to *gate (1.0 = growth).
__global__ void k_gate_targets(const float *__restrict__ rad, float *__restrict__ rt,
                               const int32_t *__restrict__ fsc, const float *__restrict__ ema,
                               int n, int frame, int warmstart, double zero_rate,
                               const int *__restrict__ zero_cnt, float fl, float fh, float gp,
                               float r_min, float r_max, float db, float *__restrict__ gate,
                               const int *__restrict__ hold_a, const int *__restrict__ hold_b) {
  // a face step whose finalize must still be finished on the host
  if ((hold_a && *hold_a) || (hold_b && *hold_b)) return;
  const double f0 = (double)*zero_cnt / (double)(n > 1 ? n : 1);
  const bool open = !(frame < warmstart || f0 > zero_rate);
  if (blockIdx.x == 0 && threadIdx.x == 0) *gate = open ? 1.0f : 0.0f;
  if (!open) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float r0 = rad[i];
  const float f = ema[i] > 0.0f ? ema[i] : (float)fsc[i];
  const float gap = 1.0f, idle = 0.003f;
  float out;
  if (f < fl - gap) {
    out = fminf(r0 * (1.0f + gp * 1.0f), r_max);
  } else if (f < fl) {
    const float g = 1.0f - smoothstep(fl - gap - db, fl, f);
    out = fminf(r0 * (1.0f + gp * g), r_max);
  } else if (f > fh + gap) {
    out = fmaxf(r0 * (1.0f - gp * 1.0f), r_min);
  } else if (f > fh) {
    const float g = smoothstep(fh, fh + gap + db, f);
    out = fmaxf(r0 * (1.0f - gp * g), r_min);
  } else {
    const float center = 0.5f * (fl + fh);
    if (f < center)
      out = fminf(r0 * (1.0f + gp * idle), r_max);
    else
      out = fmaxf(r0 * (1.0f - gp * idle), r_min);
  }
  rt[i] = out;
}

namespace fos {
int gate_targets(const float *rad, float *rad_target, const int32_t *fsc, const float *fsc_ema,
                 int n, int frame, int warmstart, double zero_rate, const int *zero_cnt,
                 int f_low, int f_high, float growth_pct, float r_min, float r_max,
                 float deadband, float *gate, hipStream_t st, const int *hold_a,
                 const int *hold_b) {
  hipLaunchKernelGGL(k_gate_targets, dim3(n > 0 ? blocks_for(n) : 1), dim3(FOS_BLOCK), 0, st,
                     rad, rad_target, fsc, fsc_ema, n, frame, warmstart, zero_rate, zero_cnt,
                     (float)f_low, (float)f_high, growth_pct, r_min, r_max, deadband, gate,
                     hold_a, hold_b);
  FOS_CHECK_LAUNCH("k_gate_targets");
  return 0;
}
}  // namespace fos

// dynamics.py:1399-1427
__device__ __forceinline__ void nudge_one(int i, float *__restrict__ rad,
                                          const float *__restrict__ rt,
                                          const float *__restrict__ lmd, float alpha, float msp,
                                          float r_min, float r_max, float back, int mode,
                                          float tol) {
  const float r0 = rad[i];
  const float t = fminf(r_max, fmaxf(r_min, rt[i]));
  float d = alpha * (t - r0);
  float b = 1.0f;
  if (mode == 1) {
    b = back;
  } else if (mode == 2) {
    b = (tol - lmd[i]) / tol;
    b = fminf(1.0f, fmaxf(0.25f, b));
  }
  d = d * b;
  const float cap = msp * r0;
  if (d > cap) d = cap;
  if (d < -cap) d = -cap;
  rad[i] = fminf(r_max, fmaxf(r_min, r0 + d));
}

__global__ void k_nudge(float *__restrict__ rad, const float *__restrict__ rt,
                        const float *__restrict__ lmd, int n, float alpha,
                        float msp, float r_min, float r_max, float back, int mode,
                        float tol) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  nudge_one(i, rad, rt, lmd, alpha, msp, r_min, r_max, back, mode, tol);
}

// run.py:733-751 and 1184-1231 with the settle counter on the device: the
// counter entering the frame is adj when the last frame skipped the gate
// (force), else the stored one, reset to adj when the last frame's device
// gate (maxd[1]) said growth (apply_gate); a positive counter nudges and
// counts down.  The counter leaving the frame goes to settle_out (a
// different buffer, so no block reads what another writes).  The global
// backpressure (mode 1) is run.py:744-749 from the overlap maximum maxd[0],
// in double as the host computes it, rounded once.
__global__ void k_settle_nudge(float *__restrict__ rad, const float *__restrict__ rt,
                               const float *__restrict__ lmd, int n, float alpha, float msp,
                               float r_min, float r_max, int mode, float tol, double tol_d,
                               const float *__restrict__ maxd, const int *__restrict__ settle_in,
                               int *__restrict__ settle_out, int adj, int force,
                               int apply_gate) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int s = force ? adj : *settle_in;
  if (apply_gate && maxd[1] > 0.5f) s = adj;
  const bool active = s > 0;
  if (i == 0) *settle_out = active ? s - 1 : s;
  if (!active || i >= n) return;
  const float back = (float)fmax(0.25, fmin(1.0, (tol_d - (double)maxd[0]) / tol_d));
  nudge_one(i, rad, rt, lmd, alpha, msp, r_min, r_max, back, mode, tol);
}

// dynamics.py:1451, int32 wraparound arithmetic
__device__ __forceinline__ int lcg_hash(int i, int frame) {
  const unsigned t = (unsigned)i + 12345u * (unsigned)frame;
  return (int)((1103515245u * t + 12345u) & 0x7fffffffu);
}

// x^y for y = fp32(1/3) = 1/3 + d, d = 2^-25 / 3, and x a positive normal
// float, when it can be decided cheaply that (float)pow((double)x, y) is the
// float returned: x^y = x^(1/3) * x^d with
//   x^(1/3) = x * u^2, u = x^(-1/3): an fp32 estimate exp2(-log2(x) / 3)
//     (relative error ~2^-17) refined by three Newton steps in double
//     u <- u + u (1 - x u^3) / 3 (quadratic: 2^-17 -> 2^-33 -> 2^-65; the
//     third covers a far worse estimate), leaving the double rounding of the
//     steps, < 2^-50 for x^(1/3);
//   x^d = exp(t) = 1 + t + t^2 / 2 with t = d ln x = d ln2 log2(x),
//     |t| < 9e-7 (t^3 / 6 < 2^-62); log2(x) = exponent + fp32 log2 of the
//     mantissa in [0.5, 1), abs. error < 2^-22, so t is off by < 2^-49.
// z is within 2^-48 of x^y.  When z (1 - 2^-44) and z (1 + 2^-44) round to
// the same float, every real within 2^-44 of z rounds to it: x^y, and the
// double pow (within a few units of 2^-53 of x^y) with it.  Otherwise
// (about one argument in 2^19) the caller evaluates the double pow.
__device__ __forceinline__ bool pow_third_fast(float x, float &out) {
  if (!(x >= 1.17549435e-38f && x <= 3.40282347e38f)) return false;
  const double xd = (double)x;
  double u = (double)__builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * -0.333333343f);
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    const double e = 1.0 - xd * (u * u * u);
    u = u + u * (e * (1.0 / 3.0));
  }
  const double c = xd * (u * u);
  int ex;
  const float m = frexpf(x, &ex);
  const double lg = (double)ex + (double)__builtin_amdgcn_logf(m);
  const double t = lg * (0x1.5555555555555p-27 * 0x1.62e42fefa39efp-1);
  const double z = c + c * (t + 0.5 * (t * t));
  const float lo = (float)(z * (1.0 - 0x1p-44));
  const float hi = (float)(z * (1.0 + 0x1p-44));
  out = lo;
  return lo == hi;
}

// fp32 power evaluated in double and rounded once: the correctly rounded fp32
// result except in ~2^-29 of cases, independent of the fp32 pow library.
// For y = fp32(1/3) (the pressure's inverse volume exponent at P = 3) the
// same float usually comes from pow_third_fast.
__device__ __forceinline__ float pow_f32(float x, float y) {
  float r;
  if (y == 0.333333343f && pow_third_fast(x, r)) return r;
  return (float)pow((double)x, (double)y);
}

// Self-check of pow_third_fast: for the positive floats with bit patterns
// first + k * stride, k < count, counts the arguments where the fast path
// answers with a float other than (float)pow((double)x, fp32(1/3)) (cnt[0])
// and those it leaves to the double pow (cnt[1]).
__global__ void __launch_bounds__(256)
    k_pow_third_check(unsigned first, unsigned count, unsigned stride,
                      unsigned long long *__restrict__ cnt) {
  unsigned long long bad = 0, slow = 0;
  for (unsigned k = blockIdx.x * 256u + threadIdx.x; k < count; k += gridDim.x * 256u) {
    const float x = __uint_as_float(first + k * stride);
    float r;
    if (pow_third_fast(x, r)) {
      const float want = (float)pow((double)x, (double)0.333333343f);
      bad += __float_as_uint(r) != __float_as_uint(want);
    } else {
      ++slow;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    bad += __shfl_xor(bad, o);
    slow += __shfl_xor(slow, o);
  }
  if ((threadIdx.x & 63) == 0) {
    if (bad) atomicAdd(&cnt[0], bad);
    if (slow) atomicAdd(&cnt[1], slow);
  }
}

// r^p for the volume: with p == 3 the product (r*r)*r in double is the
// correctly rounded cube (r*r is exact), i.e. the same value pow gives.
__device__ __forceinline__ float vol_f32(float r, float p) {
  if (p == 3.0f) {
    const double d = (double)r;
    return (float)(d * d * d);
  }
  return pow_f32(r, p);
}

// dynamics.py:1502-1543: radius change of particle i from the radii rin.
// ih: the particle's index within its foam (the rotation hash's input).
__device__ __forceinline__ float pressure_delta(int i, int ih, int n, int frame,
                                                const float *__restrict__ rin,
                                                const int32_t *__restrict__ face_ids,
                                                const int32_t *__restrict__ fsc, float k,
                                                float p_exp, float inv_p, float cap,
                                                int max_nei) {
  const int fc = fsc[i];
  float acc = 0.0f;
  if (fc > 0) {
    const float ri = rin[i];
    const float vi = vol_f32(ri, p_exp);
    const int an = fc < max_nei ? fc : max_nei;
    int idx = lcg_hash(ih, frame) % fc;
    const int32_t *row = face_ids + (int64_t)i * FOS_MAX_NEIGHBORS;
#pragma unroll 4
    for (int q = 0; q < an; ++q) {
      const int j = idx < FOS_MAX_NEIGHBORS ? row[idx] : -1;  // fsc <= 32 by construction
      idx = (idx + 1 == fc) ? 0 : idx + 1;
      if (j < 0 || j >= n || j == i) continue;
      const float vj = vol_f32(rin[j], p_exp);
      const float raw = k * (vi - vj);
      const float vmin = fminf(vi, vj);
      const float dv = fmaxf(-cap * vmin, fminf(cap * vmin, raw));
      acc = acc + (pow_f32(vi - dv, inv_p) - ri);
    }
  }
  return acc;
}

__global__ void k_pressure_delta(int n, int frame, const float *__restrict__ rad,
                                 float *__restrict__ delta_r,
                                 const int32_t *__restrict__ face_ids,
                                 const int32_t *__restrict__ fsc, float k,
                                 float p_exp, float inv_p, float cap, int max_nei) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  delta_r[i] = pressure_delta(i, i, n, frame, rad, face_ids, fsc, k, p_exp, inv_p, cap, max_nei);
}

// One Jacobi iteration (dynamics.py:1497-1557) for particle i from rin into
// rout, by a group of 16 lanes (q = lane in the group): lane q evaluates the
// q-th budgeted neighbour of the rotation, lane 0 adds the terms in rotation
// order (the serial order, bit for bit).  A budget above 16 falls back to
// the serial loop.  stats (when given) gets max|dr| and the changed count.
// Every lane of the group must call it (the shuffles span the group).
__device__ __forceinline__ void pressure_group(int i, int ih, int q, int n, int frame,
                                               const float *__restrict__ rin,
                                               float *__restrict__ rout,
                                               float *__restrict__ delta_r,
                                               const int32_t *__restrict__ face_ids,
                                               const int32_t *__restrict__ fsc, float k,
                                               float p_exp, float inv_p, float cap, int max_nei,
                                               float r_min, float r_max, float *stats) {
  const bool live = i < n;
  // the row's 32 entries (two per lane) load beside fsc[i], so the
  // neighbour's radius is the only load that waits on another
  const int32_t *row = face_ids + (int64_t)(live ? i : 0) * FOS_MAX_NEIGHBORS;
  const int e_lo = live ? row[q] : -1, e_hi = live ? row[q + 16] : -1;
  const int fc = live ? fsc[i] : 0;
  const int an = fc < max_nei ? fc : max_nei;
  const float ri = live ? rin[i] : 0.0f;
  const int idx = fc > 0 ? (int)(((unsigned)lcg_hash(ih, frame) % (unsigned)fc + (unsigned)q) %
                                 (unsigned)fc)
                         : 0;
  // every lane of the group takes part in both shuffles
  const int j_lo = __shfl(e_lo, idx & 15, 16), j_hi = __shfl(e_hi, idx & 15, 16);
  float term = 0.0f;
  bool has = false;
  if (live && fc > 0 && an <= 16 && q < an) {
    const float vi = vol_f32(ri, p_exp);
    const int j = idx < 16 ? j_lo : idx < FOS_MAX_NEIGHBORS ? j_hi : -1;
    if (!(j < 0 || j >= n || j == i)) {
      const float vj = vol_f32(rin[j], p_exp);
      const float raw = k * (vi - vj);
      const float vmin = fminf(vi, vj);
      const float dv = fmaxf(-cap * vmin, fminf(cap * vmin, raw));
      term = pow_f32(vi - dv, inv_p) - ri;
      has = true;
    }
  }
  // ordered sum over the 16-lane group
  float acc = 0.0f;
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const float t = __shfl(term, s, 16);
    const int h = __shfl((int)has, s, 16);
    if (h) acc = acc + t;
  }
  if (!live || q != 0) return;
  float d = acc;
  if (fc > 0 && an > 16)
    d = pressure_delta(i, ih, n, frame, rin, face_ids, fsc, k, p_exp, inv_p, cap, max_nei);
  delta_r[i] = d;
  if (stats) {
    const float a = fabsf(d);
    if (a > 1e-12f) {
      atomicAdd((int *)&stats[1], 1);
      atomicMax((int *)&stats[0], __float_as_int(a));
    }
  }
  rout[i] = fmaxf(r_min, fminf(r_max, ri + d));
}

// One iteration per launch, 16 particles per 256-thread block.
__global__ void __launch_bounds__(256)
    k_pressure_step(int n, int frame, const float *__restrict__ rin, float *__restrict__ rout,
                    float *__restrict__ delta_r, const int32_t *__restrict__ face_ids,
                    const int32_t *__restrict__ fsc, float k, float p_exp, float inv_p,
                    float cap, int max_nei, float r_min, float r_max, float *stats) {
  const int i = (int)(blockIdx.x * 16 + (threadIdx.x >> 4));
  pressure_group(i, i, (int)(threadIdx.x & 15), n, frame, rin, rout, delta_r, face_ids, fsc, k,
                 p_exp, inv_p, cap, max_nei, r_min, r_max, stats);
}

// ---------------------------------------------------------------------------
// All iterations of one fos_pressure_run call in one launch, with no
// grid-wide barrier.  Particle i's radius after iteration k is published
// in slot[k][i] = (tag << 32) | radius bits, written once per call (tag: the
// call's number, never 0); iteration k + 1 of i waits only on the slots k of
// its budgeted partners.  Value and tag travel in one 64-bit relaxed atomic,
// so no ordering between locations is needed.  The grid fits the device at
// once (checked against the occupancy limit before the launch) and the
// stream runs nothing beside it, so every group becomes resident and every
// wait is for a group that runs.  Each wait also ends after FLOW_SPINS polls
// (recording flags[FOS_FLAG_FLOW], which the next face step's host read
// turns into an error), so the kernel always exits.  (A cooperative launch
// would guarantee the residency too, but its dispatch costs more than the
// launches it saves.)  Slot 0 holds the input radii, so rad is read
// and written only by its own particle's group.  The arithmetic is
// pressure_group's, term for term and in the same order, so the result is
// the per-iteration launches' bit for bit.
// ---------------------------------------------------------------------------
#define FLOW_SPINS (1u << 16)

__device__ __forceinline__ void flow_put(unsigned long long *slot, unsigned tag, float r) {
  const unsigned long long v = ((unsigned long long)tag << 32) | (unsigned long long)__float_as_uint(r);
  __hip_atomic_store(slot, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long flow_peek(const unsigned long long *slot) {
  return __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The radius in `slot` once it carries `tag`; v is a value read from the
// slot earlier (polled again while its tag is an older one).
__device__ __forceinline__ float flow_get(unsigned long long v, const unsigned long long *slot,
                                          unsigned tag, int *err) {
  unsigned spins = 0;
  while ((unsigned)(v >> 32) != tag) {
    if (++spins > FLOW_SPINS) {
      atomicOr(err, 1);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
    v = flow_peek(slot);
  }
  return __uint_as_float((unsigned)v);
}

// 16 particles per 256-thread block, 16 lanes each (needs max_nei <= 16).
__global__ void __launch_bounds__(256)
    k_pressure_flow(int n, int frame, int iters, unsigned tag, float *__restrict__ rad,
                    unsigned long long *__restrict__ slot, float *__restrict__ delta_r,
                    const int32_t *__restrict__ face_ids, const int32_t *__restrict__ fsc,
                    float k, float p_exp, float inv_p, float cap, int max_nei, float r_min,
                    float r_max, float *stats, int *err) {
  const int i = (int)(blockIdx.x * 16 + (threadIdx.x >> 4));
  const int q = (int)(threadIdx.x & 15);
  const bool live = i < n;
  // the partner of lane q, fixed for the call (pressure_group's selection)
  const int32_t *row = face_ids + (int64_t)(live ? i : 0) * FOS_MAX_NEIGHBORS;
  const int e_lo = live ? row[q] : -1, e_hi = live ? row[q + 16] : -1;
  const int fc = live ? fsc[i] : 0;
  const int an = fc < max_nei ? fc : max_nei;
  const int idx = fc > 0 ? (int)(((unsigned)lcg_hash(i, frame) % (unsigned)fc + (unsigned)q) %
                                 (unsigned)fc)
                         : 0;
  const int j_lo = __shfl(e_lo, idx & 15, 16), j_hi = __shfl(e_hi, idx & 15, 16);
  int j = -1;
  if (live && fc > 0 && q < an) {
    j = idx < 16 ? j_lo : idx < FOS_MAX_NEIGHBORS ? j_hi : -1;
    if (j >= n || j == i) j = -1;
  }
  float ri = live ? rad[i] : 0.0f;
  if (live && q == 0) flow_put(slot + i, tag, ri);
  float d = 0.0f;
  unsigned long long v = j >= 0 ? flow_peek(slot + j) : 0ull;
  for (int it = 0; it < iters; ++it) {
    // a lane without a partner adds +0, which leaves the sum's bits alone
    // (the sum starts at +0 and is never -0)
    float term = 0.0f;
    if (j >= 0) {
      const float rj = flow_get(v, slot + (int64_t)it * n + j, tag, err);
      // the next iteration's read is issued before this one's arithmetic
      if (it + 1 < iters) v = flow_peek(slot + (int64_t)(it + 1) * n + j);
      const float vi = vol_f32(ri, p_exp);
      const float vj = vol_f32(rj, p_exp);
      const float raw = k * (vi - vj);
      const float vmin = fminf(vi, vj);
      const float dv = fmaxf(-cap * vmin, fminf(cap * vmin, raw));
      term = pow_f32(vi - dv, inv_p) - ri;
    }
    // ordered sum over the group's first max_nei lanes (no other lane has
    // a partner); every lane gets the same sum
    float acc = 0.0f;
    for (int s = 0; s < max_nei; ++s) acc = acc + __shfl(term, s, 16);
    d = acc;
    ri = fmaxf(r_min, fminf(r_max, ri + d));
    if (live && q == 0 && it + 1 < iters) flow_put(slot + (int64_t)(it + 1) * n + i, tag, ri);
  }
  if (!live || q != 0) return;
  delta_r[i] = d;
  if (stats) {
    const float a = fabsf(d);
    if (a > 1e-12f) {
      atomicAdd((int *)&stats[1], 1);
      atomicMax((int *)&stats[0], __float_as_int(a));
    }
  }
  rad[i] = ri;
}

// pressure_group's term for a particle of radius ri and a partner of
// radius rj (the same operations in the same order).
__device__ __forceinline__ float pressure_term(float ri, float rj, float k, float p_exp,
                                               float inv_p, float cap) {
  const float vi = vol_f32(ri, p_exp);
  const float vj = vol_f32(rj, p_exp);
  const float raw = k * (vi - vj);
  const float vmin = fminf(vi, vj);
  const float dv = fmaxf(-cap * vmin, fminf(cap * vmin, raw));
  return pow_f32(vi - dv, inv_p) - ri;
}

// One thread per particle (needs max_nei <= FLOW_MAXB), for foams too large
// for the 16-lane form to be resident at once: the budgeted
// partners are fixed for the call, so their slots are read together, their
// terms evaluated independently and summed in rotation order (the order of
// pressure_group's lanes; a missing partner adds +0, which leaves the sum's
// bits alone: the sum starts at +0 and is never -0).
#define FLOW_MAXB 16
__global__ void __launch_bounds__(256)
    k_pressure_flow_lane(int n, int frame, int iters, unsigned tag, float *__restrict__ rad,
                    unsigned long long *__restrict__ slot, float *__restrict__ delta_r,
                    const int32_t *__restrict__ face_ids, const int32_t *__restrict__ fsc,
                    float k, float p_exp, float inv_p, float cap, int max_nei, float r_min,
                    float r_max, float *stats, int *err) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  // the budget of pressure_group: partners row[(h + t) % fc], t < an
  const int fc = fsc[i];
  const int an = fc < max_nei ? fc : max_nei;
  const unsigned h = fc > 0 ? (unsigned)lcg_hash(i, frame) % (unsigned)fc : 0u;
  const int32_t *row = face_ids + (int64_t)i * FOS_MAX_NEIGHBORS;
  int pj[FLOW_MAXB];
#pragma unroll
  for (int t = 0; t < FLOW_MAXB; ++t) {
    int p = -1;
    if (t < an) {
      const unsigned ix = (h + (unsigned)t) % (unsigned)fc;
      p = ix < FOS_MAX_NEIGHBORS ? row[ix] : -1;
      if (p >= n || p == i) p = -1;
    }
    pj[t] = p;
  }
  float ri = rad[i];
  flow_put(slot + i, tag, ri);
  float d = 0.0f;
  for (int it = 0; it < iters; ++it) {
    const unsigned long long *sk = slot + (int64_t)it * n;
    unsigned long long v[FLOW_MAXB];
#pragma unroll
    for (int t = 0; t < FLOW_MAXB; ++t) v[t] = pj[t] >= 0 ? flow_peek(sk + pj[t]) : 0ull;
    float term[FLOW_MAXB];
#pragma unroll
    for (int t = 0; t < FLOW_MAXB; ++t) {
      term[t] = 0.0f;
      if (pj[t] >= 0)
        term[t] = pressure_term(ri, flow_get(v[t], sk + pj[t], tag, err), k, p_exp, inv_p, cap);
    }
    float acc = 0.0f;
#pragma unroll
    for (int t = 0; t < FLOW_MAXB; ++t) acc = acc + term[t];
    d = acc;
    ri = fmaxf(r_min, fminf(r_max, ri + d));
    if (it + 1 < iters) flow_put(slot + (int64_t)(it + 1) * n + i, tag, ri);
  }
  delta_r[i] = d;
  if (stats) {
    const float a = fabsf(d);
    if (a > 1e-12f) {
      atomicAdd((int *)&stats[1], 1);
      atomicMax((int *)&stats[0], __float_as_int(a));
    }
  }
  rad[i] = ri;
}

namespace fos {
static int g_pressure_flow = 1;
int pressure_flow_set(int on) {
  FOS_REQUIRE(on == 0 || on == 1, "pressure_flow must be 0 or 1");
  g_pressure_flow = on;
  return 0;
}
int pressure_flow_get() { return g_pressure_flow; }
}  // namespace fos

// dynamics.py:1546-1557
__global__ void k_pressure_apply(int n, float *__restrict__ rad,
                                 const float *__restrict__ delta_r, float r_min,
                                 float r_max, float *stats) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float d = delta_r[i];
  const float a = fabsf(d);
  if (a > 1e-12f) {
    atomicAdd((int *)&stats[1], 1);
    atomicMax((int *)&stats[0], __float_as_int(a));
  }
  rad[i] = fmaxf(r_min, fminf(r_max, rad[i] + d));
}

__global__ void k_minmax(const float *__restrict__ rad, int n, int *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // order-preserving int image of a float
  const int b = __float_as_int(rad[i]);
  const int key = b >= 0 ? b : (int)(0x80000000u - (unsigned)b);
  atomicMin(&out[0], key);
  atomicMax(&out[1], key);
}

// Box-Muller in double, rounded once to fp32.
__device__ __forceinline__ float box_muller(uint32_t a, uint32_t b) {
  const double u1 = (double)u01_open(a), u2 = (double)u01(b);
  return (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2));
}

// dynamics.py:1616-1634 with Philox noise keyed (seed, frame, i).
// cell_of (may be NULL): also runs the first step of the grid rebuild that
// follows (grid.py:285-309: wrap again, store, hash into a G^3 grid).
__global__ void k_brownian(float *__restrict__ pos, int n, float strength, float L,
                           float half_l, uint32_t k0, uint32_t k1, int frame,
                           int mode, int32_t *__restrict__ cell_of, int G, float inv_l) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint4 r0 = philox(make_uint4((unsigned)i, (unsigned)frame, 0u, 0u), k0, k1);
  float d[3];
  if (mode == 0) {
    d[0] = ((u01(r0.x) - 0.5f) * 2.0f) * strength;
    d[1] = ((u01(r0.y) - 0.5f) * 2.0f) * strength;
    d[2] = ((u01(r0.z) - 0.5f) * 2.0f) * strength;
  } else {
    const uint4 r1 = philox(make_uint4((unsigned)i, (unsigned)frame, 1u, 0u), k0, k1);
    d[0] = box_muller(r0.x, r0.y) * strength;
    d[1] = box_muller(r0.z, r0.w) * strength;
    d[2] = box_muller(r1.x, r1.y) * strength;
  }
  float w[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float p = pos[3 * i + a] + d[a];
    if (p < -half_l)
      p = p + L;
    else if (p >= half_l)
      p = p - L;
    w[a] = cell_of ? gwrap(p, L, inv_l) : p;
    pos[3 * i + a] = w[a];
  }
  if (cell_of)
    cell_of[i] = hash_axis(w[0], G, half_l, inv_l) * G * G + hash_axis(w[1], G, half_l, inv_l) * G +
                 hash_axis(w[2], G, half_l, inv_l);
}

// run.py:860-861 mean radius for the OU jitter, and the noise scale
// sigma * sqrt(dt) it sets (dynamics.py:626-630), into out[0] (mean) and
// out[1] (scale).  The radii are summed in double: the partial sums are
// multiples of the smallest radius's ulp below n * 2 * r_max, so they are
// exact (any order) while n * r_max / r_min < 2^28, and the mean is the
// correctly rounded double quotient rounded to fp32 -- the value
// f32(mean(rad, float64)) the host would compute.
__global__ void __launch_bounds__(1024) k_rad_mean(const float *__restrict__ rad, int n,
                                                   float dt, float jitter_rms, float tau,
                                                   float *__restrict__ out) {
  __shared__ double s[1024];
  double acc = 0.0;
  for (int i = (int)threadIdx.x; i < n; i += 1024) acc += (double)rad[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float mean_r = (float)(s[0] / (double)n);
    const float sigma = jitter_rms * mean_r * sqrtf(2.0f / tau);
    out[0] = mean_r;
    out[1] = sigma * sqrtf(dt);
  }
}

// dynamics.py:620-660; sdt_dev (when not NULL) holds the noise scale
// computed on the device and replaces sdt.
__global__ void k_ou(float *__restrict__ pos, float *__restrict__ vj,
                     const float *__restrict__ rad, int n, float tau, float dt,
                     float sdt, float capf, float L, float inv_l, uint32_t k0,
                     uint32_t k1, int frame, const float *__restrict__ sdt_dev) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (sdt_dev) sdt = *sdt_dev;
  const uint4 r = philox(make_uint4((unsigned)i, (unsigned)frame, 2u, 0u), k0, k1);
  const unsigned rr[3] = {r.x, r.y, r.z};
  float v[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float x = vj[3 * i + a];
    v[a] = x + ((-x / tau) * dt + sdt * (u01(rr[a]) - 0.5f));
  }
  const float cap = capf * (rad[i] + rad[i]);
  const float len = sqrtf(dot3(v[0], v[1], v[2])) * dt;
  if (len > cap) {
    const float s = (cap / fmaxf(1e-6f, len)) / dt;
    v[0] *= s;
    v[1] *= s;
    v[2] *= s;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    vj[3 * i + a] = v[a];
    if (pos) pos[3 * i + a] = gwrap(pos[3 * i + a] + v[a] * dt, L, inv_l);
  }
}

// dynamics.py:649-660 integrate_jitter
__global__ void k_integrate(float *__restrict__ pos, const float *__restrict__ vj, int n,
                            float dt, float L, float inv_l) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int a = 0; a < 3; ++a) pos[3 * i + a] = gwrap(pos[3 * i + a] + vj[3 * i + a] * dt, L, inv_l);
}

// dynamics.py:543-589
__global__ void k_xsph(const float *__restrict__ pos, const float *__restrict__ vel,
                       float *__restrict__ vt, int n, int G, float L, float half_l,
                       float inv_l, float lim, float eps,
                       const int32_t *__restrict__ cs, const int32_t *__restrict__ cc,
                       const int32_t *__restrict__ ci) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
  float sx = 0.0f, sy = 0.0f, sz = 0.0f;
  int cnt = 0;
  const int cx = centre_axis(px, G, L, half_l, inv_l);
  const int cy = centre_axis(py, G, L, half_l, inv_l);
  const int cz = centre_axis(pz, G, L, half_l, inv_l);
  for_stencil(cx, cy, cz, G, cs, cc, ci, [&](int j) {
    if (j == i) return;
    const float dx = gwrap(pos[3 * j] - px, L, inv_l);
    const float dy = gwrap(pos[3 * j + 1] - py, L, inv_l);
    const float dz = gwrap(pos[3 * j + 2] - pz, L, inv_l);
    if (dot3(dx, dy, dz) < lim) {
      sx += vel[3 * j];
      sy += vel[3 * j + 1];
      sz += vel[3 * j + 2];
      cnt += 1;
    }
  });
  float ax = vel[3 * i], ay = vel[3 * i + 1], az = vel[3 * i + 2];
  if (cnt > 0) {
    const float fc = (float)cnt;
    ax = sx / fc;
    ay = sy / fc;
    az = sz / fc;
  }
  vt[3 * i] = (1.0f - eps) * vel[3 * i] + eps * ax;
  vt[3 * i + 1] = (1.0f - eps) * vel[3 * i + 1] + eps * ay;
  vt[3 * i + 2] = (1.0f - eps) * vel[3 * i + 2] + eps * az;
}

__global__ void k_copy_f32(const float *__restrict__ s, float *__restrict__ d, int64_t m) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) d[i] = s[i];
}

__global__ void k_scale_f32(float *__restrict__ v, int64_t m, float f) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) v[i] *= f;
}

// topology.py:87-133
__global__ void k_topo(const float *__restrict__ pos, const float *__restrict__ rad,
                       int n, int G, float L, float half_l, float inv_l, float mult,
                       int cap, const int32_t *__restrict__ cs,
                       const int32_t *__restrict__ cc, const int32_t *__restrict__ ci,
                       int32_t *__restrict__ pairs, int32_t *__restrict__ count,
                       int32_t *__restrict__ trunc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
  const int cx = wrap_fast(centre_axis(px, G, L, half_l, inv_l), G);
  const int cy = wrap_fast(centre_axis(py, G, L, half_l, inv_l), G);
  const int cz = wrap_fast(centre_axis(pz, G, L, half_l, inv_l), G);
  int c = 0, tr = 0;
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        const int id = wrap_fast(cx + dx, G) * G * G + wrap_fast(cy + dy, G) * G + wrap_fast(cz + dz, G);
        for (int k = cs[id]; k < cs[id] + cc[id]; ++k) {
          const int j = ci[k];
          if (j <= i) continue;
          const float ddx = gwrap(pos[3 * j] - px, L, inv_l);
          const float ddy = gwrap(pos[3 * j + 1] - py, L, inv_l);
          const float ddz = gwrap(pos[3 * j + 2] - pz, L, inv_l);
          const float th = mult * (rad[i] + rad[j]);
          if (dot3(ddx, ddy, ddz) <= th * th) {
            if (c < cap) {
              pairs[(int64_t)i * cap + c] = j;
              c += 1;
            } else {
              tr = 1;
              break;
            }
          }
        }
      }
  count[i] = c;
  trunc[i] = tr;
}

// slot-record PBD pass (defined with the batched kernels below)
template <bool BATCH, int LPP>
__global__ void __launch_bounds__(256)
    k_pbd_pr(const float *__restrict__ pin, float *__restrict__ pout,
             const float4 *__restrict__ pr, const float *__restrict__ rad, int n, int G0, float L,
             float half_l, float inv_l, float gap, float mdf, float dmul,
             const int32_t *__restrict__ cs, const int32_t *__restrict__ cc,
             const int32_t *__restrict__ ci, const int32_t *__restrict__ pinst,
             const fos_inst *__restrict__ inst, const int32_t *__restrict__ npass, int pass,
             int per_xcd, int32_t *__restrict__ cell_of);
template <bool BATCH>
static void launch_pbd_pr(hipStream_t st, int n, const float *pin, float *pout, const float *pr,
                          const float *rad, int G, float L, float gap, float mdf, float dmul,
                          const int32_t *cs, const int32_t *cc, const int32_t *ci,
                          const int32_t *pinst, const fos_inst *inst, const int32_t *npass,
                          int pass, int32_t *cell_of);

extern "C" {

#define PART_ARGS_OK(n) ((n) >= 0)

int fos_pbd_pass(const float *pos_in, float *pos_out, const float *rad, int n,
                 int G, float L, float gap_factor, float max_disp_frac,
                 float disp_mult, const int32_t *cell_start,
                 const int32_t *cell_count, const int32_t *cell_indices,
                 void *stream) {
  FOS_REQUIRE(n >= 0 && G >= 3 && L > 0.0f &&
                  (n == 0 || (pos_in && pos_out && pos_in != pos_out && rad &&
                              cell_start && cell_count && cell_indices)),
              "fos_pbd_pass: bad arguments (n=%d G=%d)", n, G);
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_pbd_wave<false>, dim3((unsigned)((n + 7) / 8)), dim3(256), 0,
                     (hipStream_t)stream, pos_in, pos_out, rad, n, G, L, (float)(0.5 * (double)L),
                     (float)(1.0 / (double)L), gap_factor, max_disp_frac, disp_mult,
                     cell_start, cell_count, cell_indices, nullptr, nullptr, nullptr, 0,
                     (int32_t *)nullptr);
  FOS_CHECK_LAUNCH("k_pbd");
  return 0;
}

int fos_pbd_pass_cells(fos_ctx *ctx, const float *pos_in, float *pos_out, const float *rad,
                       int n, int G, float L, float gap_factor, float max_disp_frac,
                       float disp_mult, const int32_t *cell_start, const int32_t *cell_count,
                       const int32_t *cell_indices, const float *cell_pr, void *stream) {
  FOS_REQUIRE(ctx && n >= 0 && n <= ctx->max_n && G >= 3 &&
                  (int64_t)G * G * G <= ctx->max_cells && L > 0.0f &&
                  (n == 0 || (pos_in && pos_out && pos_in != pos_out && rad && cell_start &&
                              cell_count && cell_indices)),
              "fos_pbd_pass_cells: bad arguments (n=%d G=%d)", n, G);
  if (n == 0) return 0;
  if (cell_pr) {
    launch_pbd_pr<false>((hipStream_t)stream, n, pos_in, pos_out, cell_pr, rad, G, L, gap_factor,
                         max_disp_frac, disp_mult, cell_start, cell_count, cell_indices, nullptr,
                         nullptr, nullptr, 0, ctx->cell_of);
    FOS_CHECK_LAUNCH("k_pbd_pr");
    return 0;
  }
  hipLaunchKernelGGL(k_pbd_wave<false>, dim3((unsigned)((n + 7) / 8)), dim3(256), 0,
                     (hipStream_t)stream, pos_in, pos_out, rad, n, G, L, (float)(0.5 * (double)L),
                     (float)(1.0 / (double)L), gap_factor, max_disp_frac, disp_mult,
                     cell_start, cell_count, cell_indices, nullptr, nullptr, nullptr, 0,
                     ctx->cell_of);
  FOS_CHECK_LAUNCH("k_pbd");
  return 0;
}

int fos_max_overlap(const float *pos, const float *rad, int n, int G, float L,
                    const int32_t *cell_start, const int32_t *cell_count,
                    const int32_t *cell_indices, float *local_max_depth,
                    float *max_depth, void *stream) {
  FOS_REQUIRE(n >= 0 && G >= 3 && L > 0.0f && max_depth &&
                  (n == 0 || (pos && rad && cell_start && cell_count &&
                              cell_indices && local_max_depth)),
              "fos_max_overlap: bad arguments (n=%d G=%d)", n, G);
  int rc = fos_fill_f32(max_depth, 1, 0.0f, stream);
  if (rc || n == 0) return rc;
  hipLaunchKernelGGL(k_max_overlap_wave<false>, dim3((unsigned)((n + 7) / 8)), dim3(256), 0,
                     (hipStream_t)stream, pos, rad, n, G, L, (float)(0.5 * (double)L),
                     (float)(1.0 / (double)L), cell_start, cell_count, cell_indices,
                     local_max_depth, max_depth, nullptr, nullptr);
  FOS_CHECK_LAUNCH("k_max_overlap");
  return 0;
}

int fos_fsc_targets(const float *rad, float *rad_target, const int32_t *fsc,
                    const float *fsc_ema, int n, int f_low, int f_high,
                    float growth_pct, float r_min, float r_max, float deadband,
                    void *stream) {
  FOS_REQUIRE(n >= 0 && (n == 0 || (rad && rad_target && fsc && fsc_ema)),
              "fos_fsc_targets: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_fsc_targets, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream, rad, rad_target, fsc, fsc_ema, n,
                     (float)f_low, (float)f_high, growth_pct, r_min, r_max, deadband,
                     (const float *)nullptr);
  FOS_CHECK_LAUNCH("k_fsc_targets");
  return 0;
}

int fos_fsc_targets_gated(const float *rad, float *rad_target, const int32_t *fsc,
                          const float *fsc_ema, int n, int f_low, int f_high, float growth_pct,
                          float r_min, float r_max, float deadband, const float *gate,
                          void *stream) {
  FOS_REQUIRE(n >= 0 && gate && (n == 0 || (rad && rad_target && fsc && fsc_ema)),
              "fos_fsc_targets_gated: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_fsc_targets, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream, rad, rad_target, fsc, fsc_ema, n,
                     (float)f_low, (float)f_high, growth_pct, r_min, r_max, deadband, gate);
  FOS_CHECK_LAUNCH("k_fsc_targets");
  return 0;
}

int fos_fsc_gate(fos_ctx *ctx, const int32_t *fsc, int n, int frame, int warmstart_frames,
                 double fsc_zero_rate, float *gate, void *stream) {
  FOS_REQUIRE(ctx && gate && n >= 0 && (n == 0 || fsc), "fos_fsc_gate: bad arguments");
  // counters: ctx->flags[4..5], zero between calls
  hipLaunchKernelGGL(k_count_gate, dim3(n > 0 ? blocks_for(n) : 1), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream, fsc, n, frame, warmstart_frames, fsc_zero_rate,
                     ctx->flags + 4, gate);
  FOS_CHECK_LAUNCH("k_count_gate");
  return 0;
}

int fos_controller_stats(fos_ctx *ctx, const float *rad, const float *rad_target,
                         const float *fsc_ema, int n, int f_low, int f_high,
                         int64_t *out_i, double *out_f, void *stream) {
  FOS_REQUIRE(ctx && n >= 0 && out_i && out_f &&
                  (n == 0 || (rad && rad_target && fsc_ema)),
              "fos_controller_stats: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  unsigned long long *cnt = (unsigned long long *)ctx->iscratch;
  int *mx = ctx->iscratch + 6;
  double *sum = (double *)ctx->fscratch;
  int rc = fos_fill_i32(ctx->iscratch, 8, 0, stream);
  if (rc) return rc;
  rc = fos_fill_i32((int32_t *)sum, 2, 0, stream);
  if (rc) return rc;
  if (n > 0) {
    hipLaunchKernelGGL(k_controller_stats, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, st,
                       rad, rad_target, fsc_ema, n, (float)f_low, (float)f_high, cnt,
                       sum, mx);
    FOS_CHECK_LAUNCH("k_controller_stats");
  }
  unsigned long long hc[3];
  int hm = 0;
  double hs = 0.0;
  FOS_HIP(hipMemcpyAsync(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, st));
  FOS_HIP(hipMemcpyAsync(&hm, mx, sizeof(hm), hipMemcpyDeviceToHost, st));
  FOS_HIP(hipMemcpyAsync(&hs, sum, sizeof(hs), hipMemcpyDeviceToHost, st));
  FOS_HIP(hipStreamSynchronize(st));
  out_i[0] = (int64_t)hc[0];
  out_i[1] = (int64_t)hc[1];
  out_i[2] = (int64_t)hc[2];
  float m;
  memcpy(&m, &hm, sizeof(m));
  out_f[0] = hs / (double)(n > 1 ? n : 1);
  out_f[1] = (double)m;
  return 0;
}

int fos_nudge_radii(float *rad, const float *rad_target,
                    const float *local_max_depth, int n, float alpha,
                    float max_step_pct, float r_min, float r_max, float back_global,
                    int mode, float contact_tol, void *stream) {
  FOS_REQUIRE(n >= 0 && (n == 0 || (rad && rad_target && local_max_depth)),
              "fos_nudge_radii: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_nudge, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream, rad, rad_target, local_max_depth, n, alpha,
                     max_step_pct, r_min, r_max, back_global, mode, contact_tol);
  FOS_CHECK_LAUNCH("k_nudge");
  return 0;
}

int fos_settle_nudge(float *rad, const float *rad_target, const float *local_max_depth, int n,
                     float alpha, float max_step_pct, float r_min, float r_max, int mode,
                     float contact_tol, double contact_tol_d, const float *maxd,
                     const int32_t *settle_in, int32_t *settle_out, int adjustment_frames,
                     int force, int apply_gate, void *stream) {
  FOS_REQUIRE(n >= 0 && maxd && settle_in && settle_out && settle_in != settle_out &&
                  contact_tol_d != 0.0 && (n == 0 || (rad && rad_target && local_max_depth)),
              "fos_settle_nudge: bad arguments");
  // one thread at least, to carry the counter
  hipLaunchKernelGGL(k_settle_nudge, dim3(blocks_for(n > 0 ? n : 1)), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream, rad, rad_target, local_max_depth, n, alpha,
                     max_step_pct, r_min, r_max, mode, contact_tol, contact_tol_d, maxd,
                     (const int *)settle_in, (int *)settle_out, adjustment_frames, force,
                     apply_gate);
  FOS_CHECK_LAUNCH("k_settle_nudge");
  return 0;
}

int fos_pressure_iter(int n, int frame, float *rad, float *delta_r,
                      const int32_t *face_ids, const int32_t *fsc, float k,
                      float p_exp, float pair_cap, int max_nei, float r_min,
                      float r_max, float *stats, void *stream) {
  FOS_REQUIRE(n >= 0 && stats && (n == 0 || (rad && delta_r && face_ids && fsc)),
              "fos_pressure_iter: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  int rc = fos_fill_f32(stats, 2, 0.0f, stream);
  if (rc || n == 0) return rc;
  hipLaunchKernelGGL(k_pressure_delta, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, st, n,
                     frame, rad, delta_r, face_ids, fsc, k, p_exp, 1.0f / p_exp,
                     pair_cap, max_nei);
  FOS_CHECK_LAUNCH("k_pressure_delta");
  hipLaunchKernelGGL(k_pressure_apply, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, st, n,
                     rad, delta_r, r_min, r_max, stats);
  FOS_CHECK_LAUNCH("k_pressure_apply");
  return 0;
}

// Blocks of a flow kernel the device holds at once (asked once per context;
// -1 when the query fails).
static int flow_capacity(fos_ctx *ctx, int &cached, const void *kernel) {
  if (cached == 0) {
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) !=
            hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess ||
        per_cu < 1)
      cached = -1;
    else
      cached = per_cu * cus;
  }
  return cached;
}

int fos_pressure_run(fos_ctx *ctx, int n, int frame, int iters, float *rad,
                     float *delta_r, const int32_t *face_ids, const int32_t *fsc, float k,
                     float p_exp, float pair_cap, int max_nei, float r_min, float r_max,
                     float *stats, void *stream) {
  FOS_REQUIRE(ctx && n >= 0 && n <= ctx->max_n && iters >= 0 &&
                  (n == 0 || (rad && delta_r && face_ids && fsc)),
              "fos_pressure_run: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  if (stats) {
    int rc = fos_fill_f32(stats, 2, 0.0f, stream);
    if (rc) return rc;
  }
  if (n == 0 || iters == 0) return 0;
  // all iterations in one launch when every group fits on the chip at once
  // (not under graph capture: the tag is a host counter)
  // (16 lanes per particle where the grid fits the device at once, else one
  // lane per particle: measured faster respectively at N = 10,000 and 50,000)
  hipStreamCaptureStatus cap_st = hipStreamCaptureStatusNone;
  FOS_HIP(hipStreamIsCapturing(st, &cap_st));
  const unsigned grid16 = (unsigned)((n + 15) / 16), grid1 = (unsigned)((n + 255) / 256);
  const bool flow_ok = g_pressure_flow && iters >= 2 && iters <= FOS_FLOW_MAX_ITERS &&
                       max_nei <= 16 && n <= ctx->flow_n && cap_st == hipStreamCaptureStatusNone;
  const bool use16 =
      flow_ok && (int)grid16 <= flow_capacity(ctx, ctx->flow_blocks, (const void *)k_pressure_flow);
  const bool use1 =
      flow_ok && !use16 &&
      (int)grid1 <= flow_capacity(ctx, ctx->flow_blocks_lane, (const void *)k_pressure_flow_lane);
  if (use16 || use1) {
    unsigned tag = ++ctx->flow_tag;
    if (tag == 0) tag = ++ctx->flow_tag;
    float inv_p = 1.0f / p_exp;
    float *rad_a = rad;
    const int32_t *fids = face_ids, *fs = fsc;
    int *err = ctx->flags + FOS_FLAG_FLOW;
    unsigned long long *slots = ctx->flow;
    if (use16)
      hipLaunchKernelGGL(k_pressure_flow, dim3(grid16), dim3(256), 0, st, n, frame, iters, tag,
                         rad_a, slots, delta_r, fids, fs, k, p_exp, inv_p, pair_cap, max_nei,
                         r_min, r_max, stats, err);
    else
      hipLaunchKernelGGL(k_pressure_flow_lane, dim3(grid1), dim3(256), 0, st, n, frame, iters,
                         tag, rad_a, slots, delta_r, fids, fs, k, p_exp, inv_p, pair_cap, max_nei,
                         r_min, r_max, stats, err);
    FOS_CHECK_LAUNCH("k_pressure_flow");
    return 0;
  }
  float *a = rad, *b = ctx->rad_tmp;
  for (int it = 0; it < iters; ++it) {
    hipLaunchKernelGGL(k_pressure_step, dim3((unsigned)((n + 15) / 16)), dim3(256), 0, st, n, frame,
                       a, b, delta_r, face_ids, fsc, k, p_exp, 1.0f / p_exp, pair_cap,
                       max_nei, r_min, r_max, it == iters - 1 ? stats : nullptr);
    FOS_CHECK_LAUNCH("k_pressure_step");
    float *t = a;
    a = b;
    b = t;
  }
  if (a != rad) {
    hipLaunchKernelGGL(k_copy_f32, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, st, a, rad,
                       (int64_t)n);
    FOS_CHECK_LAUNCH("k_copy_f32");
  }
  return 0;
}

int fos_pressure_stats(fos_ctx *ctx, const float *rad, int n, float *out,
                       void *stream) {
  FOS_REQUIRE(ctx && n >= 1 && rad && out, "fos_pressure_stats: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  int *mm = ctx->iscratch;
  int rc = fos_fill_i32(mm, 1, 0x7fffffff, stream);
  if (rc) return rc;
  rc = fos_fill_i32(mm + 1, 1, (int32_t)0x80000000, stream);
  if (rc) return rc;
  hipLaunchKernelGGL(k_minmax, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, st, rad, n, mm);
  FOS_CHECK_LAUNCH("k_minmax");
  int h[2];
  FOS_HIP(hipMemcpyAsync(h, mm, sizeof(h), hipMemcpyDeviceToHost, st));
  FOS_HIP(hipStreamSynchronize(st));
  for (int q = 0; q < 2; ++q) {
    const int key = h[q];
    const int b = key >= 0 ? key : (int)(0x80000000u - (unsigned)key);
    memcpy(&out[q], &b, sizeof(float));
  }
  return 0;
}

int fos_pow_third_check(unsigned first, unsigned count, unsigned stride,
                        unsigned long long *counts, void *stream) {
  FOS_REQUIRE(counts && stride >= 1 && (unsigned long long)first +
                                               (unsigned long long)(count ? count - 1 : 0) * stride <
                                           0x80000000ull,
              "fos_pow_third_check: the bit patterns must stay below 0x80000000");
  if (count == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  const unsigned blocks = count / 256 >= 8192 ? 8192u : (count + 255) / 256;
  hipLaunchKernelGGL(k_pow_third_check, dim3(blocks), dim3(256), 0, st, first, count, stride,
                     counts);
  FOS_CHECK_LAUNCH("k_pow_third_check");
  return 0;
}

int fos_brownian(float *pos, int n, float strength, float L, uint64_t seed,
                 int frame, int mode, void *stream) {
  FOS_REQUIRE(n >= 0 && (mode == 0 || mode == 1) && L > 0.0f && (n == 0 || pos),
              "fos_brownian: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_brownian, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream, pos, n, strength, L, L * 0.5f,
                     (uint32_t)seed, (uint32_t)(seed >> 32), frame, mode, (int32_t *)nullptr, 0,
                     0.0f);
  FOS_CHECK_LAUNCH("k_brownian");
  return 0;
}

int fos_brownian_cells(fos_ctx *ctx, float *pos, int n, float strength, float L, uint64_t seed,
                       int frame, int mode, int G, void *stream) {
  FOS_REQUIRE(ctx && n >= 0 && n <= ctx->max_n && (mode == 0 || mode == 1) && L > 0.0f &&
                  G >= 3 && G <= 1290 && (int64_t)G * G * G <= ctx->max_cells && (n == 0 || pos),
              "fos_brownian_cells: bad arguments (n=%d G=%d)", n, G);
  if (n == 0) return 0;
  // half_l = L/2 is exact in fp32, so the noise wrap and the hash share it
  hipLaunchKernelGGL(k_brownian, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream, pos, n, strength, L, L * 0.5f,
                     (uint32_t)seed, (uint32_t)(seed >> 32), frame, mode, ctx->cell_of, G,
                     (float)(1.0 / (double)L));
  FOS_CHECK_LAUNCH("k_brownian");
  return 0;
}

int fos_ou_jitter(float *pos, float *v_jit, const float *rad, int n, float mean_r,
                  float dt, float jitter_rms, float jitter_tau,
                  float drift_cap_frac, float L, uint64_t seed, int frame,
                  void *stream) {
  FOS_REQUIRE(n >= 0 && L > 0.0f && dt > 0.0f && (n == 0 || (v_jit && rad)),
              "fos_ou_jitter: bad arguments");
  if (n == 0) return 0;
  const float tau = jitter_tau > 1e-6f ? jitter_tau : 1e-6f;
  const float sigma = jitter_rms * mean_r * sqrtf(2.0f / tau);
  const float sdt = sigma * sqrtf(dt);
  hipLaunchKernelGGL(k_ou, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, (hipStream_t)stream,
                     pos, v_jit, rad, n, tau, dt, sdt, drift_cap_frac, L,
                     (float)(1.0 / (double)L), (uint32_t)seed ^ 0x5bd1e995u,
                     (uint32_t)(seed >> 32), frame, (const float *)nullptr);
  FOS_CHECK_LAUNCH("k_ou");
  return 0;
}

int fos_ou_jitter_mean(fos_ctx *ctx, float *pos, float *v_jit, const float *rad, int n,
                       float dt, float jitter_rms, float jitter_tau, float drift_cap_frac,
                       float L, uint64_t seed, int frame, float *mean_out, void *stream) {
  FOS_REQUIRE(ctx && n >= 0 && L > 0.0f && dt > 0.0f && (n == 0 || (v_jit && rad)),
              "fos_ou_jitter_mean: bad arguments");
  if (n == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  const float tau = jitter_tau > 1e-6f ? jitter_tau : 1e-6f;
  float *m = ctx->fscratch + 8;  // [0] mean radius, [1] sigma * sqrt(dt)
  hipLaunchKernelGGL(k_rad_mean, dim3(1), dim3(1024), 0, st, rad, n, dt, jitter_rms, tau, m);
  FOS_CHECK_LAUNCH("k_rad_mean");
  hipLaunchKernelGGL(k_ou, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, st, pos, v_jit, rad, n, tau,
                     dt, 0.0f, drift_cap_frac, L, (float)(1.0 / (double)L),
                     (uint32_t)seed ^ 0x5bd1e995u, (uint32_t)(seed >> 32), frame,
                     (const float *)(m + 1));
  FOS_CHECK_LAUNCH("k_ou");
  if (mean_out) FOS_HIP(hipMemcpyAsync(mean_out, m, sizeof(float), hipMemcpyDeviceToDevice, st));
  return 0;
}

int fos_integrate_jitter(float *pos, const float *v_jit, int n, float dt, float L,
                         void *stream) {
  FOS_REQUIRE(n >= 0 && L > 0.0f && (n == 0 || (pos && v_jit)),
              "fos_integrate_jitter: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_integrate, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, (hipStream_t)stream,
                     pos, v_jit, n, dt, L, (float)(1.0 / (double)L));
  FOS_CHECK_LAUNCH("k_integrate");
  return 0;
}

int fos_xsph(const float *pos, float *vel, float *vel_tmp, int n, int G, float L,
             float lim_sq, float eps, const int32_t *cell_start,
             const int32_t *cell_count, const int32_t *cell_indices, void *stream) {
  FOS_REQUIRE(n >= 0 && G >= 3 && L > 0.0f &&
                  (n == 0 || (pos && vel && vel_tmp && cell_start && cell_count &&
                              cell_indices)),
              "fos_xsph: bad arguments");
  if (n == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_xsph, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, st, pos, vel,
                     vel_tmp, n, G, L, (float)(0.5 * (double)L), (float)(1.0 / (double)L),
                     lim_sq, eps, cell_start, cell_count, cell_indices);
  FOS_CHECK_LAUNCH("k_xsph");
  hipLaunchKernelGGL(k_copy_f32, dim3(blocks_for((int64_t)n * 3)), dim3(FOS_BLOCK), 0, st,
                     vel_tmp, vel, (int64_t)n * 3);
  FOS_CHECK_LAUNCH("k_copy_f32");
  return 0;
}

int fos_damp(float *vel, int n, float factor, void *stream) {
  FOS_REQUIRE(n >= 0 && (n == 0 || vel), "fos_damp: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_scale_f32, dim3(blocks_for((int64_t)n * 3)), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream, vel, (int64_t)n * 3, factor);
  FOS_CHECK_LAUNCH("k_scale_f32");
  return 0;
}

int fos_topo_candidates(const float *pos, const float *rad, int n, int G, float L,
                        float mult, int cap, const int32_t *cell_start,
                        const int32_t *cell_count, const int32_t *cell_indices,
                        int32_t *pairs, int32_t *count, int32_t *truncated,
                        void *stream) {
  FOS_REQUIRE(n >= 0 && G >= 3 && cap >= 1 && L > 0.0f &&
                  (n == 0 || (pos && rad && cell_start && cell_count && cell_indices &&
                              pairs && count && truncated)),
              "fos_topo_candidates: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_topo, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, (hipStream_t)stream,
                     pos, rad, n, G, L, (float)(0.5 * (double)L), (float)(1.0 / (double)L),
                     mult, cap, cell_start, cell_count, cell_indices, pairs, count,
                     truncated);
  FOS_CHECK_LAUNCH("k_topo");
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches of foams (no reference counterpart): the per-particle kernels above
// with each particle's foam parameters (radius bounds, FSC band, seed, index
// within the foam) read from the foam descriptors.
// ---------------------------------------------------------------------------

__global__ void k_nudge_b(float *__restrict__ rad, const float *__restrict__ rt,
                          const float *__restrict__ lmd, const int32_t *__restrict__ pinst,
                          const fos_inst *__restrict__ inst, const int32_t *__restrict__ active,
                          const float *__restrict__ back, int n, float alpha, float msp,
                          int mode, float tol) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = pinst[i];
  if (!active[b]) return;
  const float r_min = inst[b].r_min, r_max = inst[b].r_max;
  const float r0 = rad[i];
  const float t = fminf(r_max, fmaxf(r_min, rt[i]));
  float d = alpha * (t - r0);
  float f = 1.0f;
  if (mode == 1) {
    f = back[b];
  } else if (mode == 2) {
    f = (tol - lmd[i]) / tol;
    f = fminf(1.0f, fmaxf(0.25f, f));
  }
  d = d * f;
  const float cap = msp * r0;
  if (d > cap) d = cap;
  if (d < -cap) d = -cap;
  rad[i] = fminf(r_max, fmaxf(r_min, r0 + d));
}

__global__ void k_fsc_targets_b(const float *__restrict__ rad, float *__restrict__ rt,
                                const int32_t *__restrict__ fsc, const float *__restrict__ ema,
                                const int32_t *__restrict__ pinst,
                                const fos_inst *__restrict__ inst,
                                const int32_t *__restrict__ active, int n, float gp, float db) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = pinst[i];
  if (!active[b]) return;
  const float fl = inst[b].f_low, fh = inst[b].f_high;
  const float r_min = inst[b].r_min, r_max = inst[b].r_max;
  const float r0 = rad[i];
  const float f = ema[i] > 0.0f ? ema[i] : (float)fsc[i];
  const float gap = 1.0f, idle = 0.003f;
  float out;
  if (f < fl - gap) {
    out = fminf(r0 * (1.0f + gp * 1.0f), r_max);
  } else if (f < fl) {
    const float g = 1.0f - smoothstep(fl - gap - db, fl, f);
    out = fminf(r0 * (1.0f + gp * g), r_max);
  } else if (f > fh + gap) {
    out = fmaxf(r0 * (1.0f - gp * 1.0f), r_min);
  } else if (f > fh) {
    const float g = smoothstep(fh, fh + gap + db, f);
    out = fmaxf(r0 * (1.0f - gp * g), r_min);
  } else {
    const float center = 0.5f * (fl + fh);
    if (f < center)
      out = fminf(r0 * (1.0f + gp * idle), r_max);
    else
      out = fmaxf(r0 * (1.0f - gp * idle), r_min);
  }
  rt[i] = out;
}

// frame_dev (may be NULL): the frame number in device memory (a captured
// frame reads it at replay), else frame.
__global__ void __launch_bounds__(256)
    k_pressure_step_b(int n, int frame, const int32_t *__restrict__ frame_dev,
                      const float *__restrict__ rin, float *__restrict__ rout,
                      float *__restrict__ delta_r, const int32_t *__restrict__ face_ids,
                      const int32_t *__restrict__ fsc, const int32_t *__restrict__ pinst,
                      const fos_inst *__restrict__ inst, float k, float p_exp, float inv_p,
                      float cap, int max_nei) {
  if (frame_dev) frame = *frame_dev;
  const int i = (int)(blockIdx.x * 16 + (threadIdx.x >> 4));
  int ih = i;
  float r_min = 0.0f, r_max = 0.0f;
  if (i < n) {
    const int b = pinst[i];
    ih = i - inst[b].p0;
    r_min = inst[b].r_min;
    r_max = inst[b].r_max;
  }
  pressure_group(i, ih, (int)(threadIdx.x & 15), n, frame, rin, rout, delta_r, face_ids, fsc, k,
                 p_exp, inv_p, cap, max_nei, r_min, r_max, nullptr);
}

__global__ void k_brownian_b(float *__restrict__ pos, const int32_t *__restrict__ pinst,
                             const fos_inst *__restrict__ inst, int n, float strength, float L,
                             float half_l, int frame, const int32_t *__restrict__ frame_dev,
                             int mode) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  if (frame_dev) frame = *frame_dev;
  const int b = pinst[g];
  const unsigned i = (unsigned)(g - inst[b].p0);
  const uint32_t k0 = inst[b].seed, k1 = 0u;
  const uint4 r0 = philox(make_uint4(i, (unsigned)frame, 0u, 0u), k0, k1);
  float d[3];
  if (mode == 0) {
    d[0] = ((u01(r0.x) - 0.5f) * 2.0f) * strength;
    d[1] = ((u01(r0.y) - 0.5f) * 2.0f) * strength;
    d[2] = ((u01(r0.z) - 0.5f) * 2.0f) * strength;
  } else {
    const uint4 r1 = philox(make_uint4(i, (unsigned)frame, 1u, 0u), k0, k1);
    d[0] = box_muller(r0.x, r0.y) * strength;
    d[1] = box_muller(r0.z, r0.w) * strength;
    d[2] = box_muller(r1.x, r1.y) * strength;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float p = pos[3 * g + a] + d[a];
    if (p < -half_l)
      p = p + L;
    else if (p >= half_l)
      p = p - L;
    pos[3 * g + a] = p;
  }
}

// out[b] += number of particles of foam b with fsc == 0 (out zeroed first).
__global__ void k_count_zero_b(const int32_t *__restrict__ fsc, const int32_t *__restrict__ pinst,
                               int n, int32_t *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < n;
  const int b = live ? pinst[i] : -1;
  const bool z = live && fsc[i] == 0;
  // one atomic per wave when its particles share a foam
  const int b0 = __shfl(b, 0, 64);
  const unsigned long long same = __ballot(!live || b == b0);
  if (same == ~0ull) {
    const int c = __popcll(__ballot(z));
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(&out[b0], c);
  } else if (z) {
    atomicAdd(&out[b], 1);
  }
}

// The batched PBD pass reads each stencil cell's occupants from the
// cell-ordered (position, radius) records of the grid rebuild, so a lane's
// loads are contiguous and independent of the id load.  Lanes first count
// their cell's corrections, a prefix over the 27 lanes gives each its place
// in the serial (cell, occupant) order, and the lanes then write the
// corrections there for lane 0 to add in that order, PBD_HITS at a time
// (32 in the 8-lane form, whose 32 particles per block would otherwise hold
// 18 KB of windows: with 16 KB of LDS and at most 64 VGPRs it keeps eight
// waves per SIMD; the batched C4 frame +0.5%).
#define PBD_HITS 48

// pbd_hit's contact test alone (dist < ts with dist = sqrtf(d.d + 1e-12)),
// rejecting first, without the square root, every pair whose squared
// distance s exceeds fl(fl(ts*ts) * 1.0001f) >= 1.0000999 ts^2: then sqrt(s) >
// 1.00004 ts, and the correctly rounded sqrtf(s) is still above ts, so the
// exact test would reject it too.
// gwrap; with U (L == 1) the products by L and 1/L are exact, so d -
// floorf(d + 0.5f) gives the same bits in two fewer operations.
template <bool U>
__device__ __forceinline__ float gw(float d, float L, float inv_l) {
  return U ? d - floorf(d + 0.5f) : gwrap(d, L, inv_l);
}

template <bool U>
__device__ __forceinline__ bool pbd_contact(float qx, float qy, float qz, float rj, float px,
                                            float py, float pz, float ri, float gap, float L,
                                            float inv_l) {
  const float dx = gw<U>(qx - px, L, inv_l);
  const float dy = gw<U>(qy - py, L, inv_l);
  const float dz = gw<U>(qz - pz, L, inv_l);
  const float s = dot3(dx, dy, dz) + 1e-12f;
  const float ts = (ri + rj) * gap;
  if (s > ts * ts * 1.0001f) return false;
  return sqrtf(s) < ts;
}

template <bool U = false>
__device__ __forceinline__ bool pbd_hit(float qx, float qy, float qz, float rj, float px,
                                        float py, float pz, float ri, float maxd, float gap,
                                        float L, float inv_l, float &vx, float &vy, float &vz) {
  const float dx = gw<U>(qx - px, L, inv_l);
  const float dy = gw<U>(qy - py, L, inv_l);
  const float dz = gw<U>(qz - pz, L, inv_l);
  const float dist = sqrtf(dot3(dx, dy, dz) + 1e-12f);
  const float ts = (ri + rj) * gap;
  if (!(dist < ts)) return false;
  const float overlap = ts - dist;
  const float ux = dx / dist, uy = dy / dist, uz = dz / dist;
  const float mag = fminf(overlap * 0.5f, maxd);
  vx = fmaxf(-maxd, fminf(maxd, -ux * mag));
  vy = fmaxf(-maxd, fminf(maxd, -uy * mag));
  vz = fmaxf(-maxd, fminf(maxd, -uz * mag));
  return true;
}

// LPP lanes per particle (8, 16 or 32; 256 / LPP particles per block): lane
// l walks stencil cells [l*CPL, l*CPL + CPL) in order, so the hits, placed by
// an exclusive prefix over the lanes, stay in stencil-then-slot order.  The
// walk records a lane's first PBD_HK hit slots in LDS, so a lane with at
// most that many hits writes its corrections from those slots instead of
// walking its cells again.  U: L == 1.
#define PBD_HK 4
template <bool BATCH, int LPP, bool U = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
    k_pbd_pr(const float *__restrict__ pin, float *__restrict__ pout,
             const float4 *__restrict__ pr, const float *__restrict__ rad, int n, int G0, float L,
             float half_l, float inv_l, float gap, float mdf, float dmul,
             const int32_t *__restrict__ cs, const int32_t *__restrict__ cc,
             const int32_t *__restrict__ ci, const int32_t *__restrict__ pinst,
             const fos_inst *__restrict__ inst, const int32_t *__restrict__ npass, int pass,
             int per_xcd, int32_t *__restrict__ cell_of) {
  constexpr int NPB = 256 / LPP, CPL = (27 + LPP - 1) / LPP;
  constexpr int HITS = LPP == 8 ? 32 : PBD_HITS;
  __shared__ float s_v[NPB][HITS][3];
  __shared__ int s_hk[PBD_HK][256];
  const int t = (int)threadIdx.x;
  const int grp = (int)(threadIdx.x / LPP), l = (int)(threadIdx.x % LPP);
  // particles in cell order (slot k of the grid holds particle ci[k]): a
  // workgroup's particles and an XCD's run of workgroups read the same
  // neighbourhood of slot records from their L2
  const int kk = xcd_tile((int)blockIdx.x, per_xcd) * NPB + grp;
  const int i = kk < n ? ci[kk] : n;
  bool live = i < n;
  int G = G0;
  if (BATCH && live) {
    const int b = pinst[i];
    if (pass >= npass[b]) {
      if (l < 3) pout[3 * i + l] = pin[3 * i + l];
      live = false;
    } else {
      G = inst[b].G;
      cs += inst[b].c0;
      cc += inst[b].c0;
    }
  }
  float px = 0.f, py = 0.f, pz = 0.f, ri = 0.f, maxd = 0.f;
  int cx = 0, cy = 0, cz = 0;
  int st[CPL], ct[CPL];
#pragma unroll
  for (int u = 0; u < CPL; ++u) st[u] = ct[u] = 0;
  if (live) {
    px = pin[3 * i];
    py = pin[3 * i + 1];
    pz = pin[3 * i + 2];
    ri = rad[i];
    maxd = mdf * ri;
    cx = centre_axis(px, G, L, half_l, inv_l);
    cy = centre_axis(py, G, L, half_l, inv_l);
    cz = centre_axis(pz, G, L, half_l, inv_l);
#pragma unroll
    for (int u = 0; u < CPL; ++u) {
      const int c = l * CPL + u;
      if (c < 27) {
        const int id = wrap_fast(cx + c / 9 - 1, G) * G * G +
                       wrap_fast(cy + (c / 3) % 3 - 1, G) * G + wrap_fast(cz + c % 3 - 1, G);
        st[u] = cs[id];
        ct[u] = cc[id];
      }
    }
  }
  // particle i sits in slot kk, so "slot k is not i" is k != kk
  int h = 0;
  float vx, vy, vz;
#pragma unroll
  for (int u = 0; u < CPL; ++u)
    for (int k = st[u]; k < st[u] + ct[u]; ++k) {
      const float4 q = pr[k];
      if (k != kk && pbd_contact<U>(q.x, q.y, q.z, q.w, px, py, pz, ri, gap, L, inv_l)) {
        if (h < PBD_HK) s_hk[h][t] = k;
        ++h;
      }
    }
  int x = h;  // inclusive prefix over the particle's lanes
  for (int o = 1; o < LPP; o <<= 1) {
    const int y = __shfl_up(x, o, LPP);
    if (l >= o) x += y;
  }
  const int H = __shfl(x, LPP - 1, LPP);
  // The hits go through LDS in windows of PBD_HITS, lane 0 adding each
  // window in order, so a particle with more hits than one window (clustered
  // foams) keeps its lanes instead of walking the 27 cells alone; every
  // group of the block runs the block's largest window count (one in the
  // usual case).
  __shared__ int s_rounds;
  if (threadIdx.x == 0) s_rounds = 1;
  __syncthreads();
  if (live && l == 0 && H > HITS) atomicMax(&s_rounds, (H + HITS - 1) / HITS);
  __syncthreads();
  const int rounds = s_rounds;
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
  for (int r = 0; r < rounds; ++r) {
    const int lo = r * HITS, hi = lo + HITS;
    if (live && h > 0 && x > lo && x - h < hi) {
      int m = x - h;
      if (h <= PBD_HK) {
        for (int j = 0; j < h; ++j, ++m) {
          const float4 q = pr[s_hk[j][t]];
          pbd_hit<U>(q.x, q.y, q.z, q.w, px, py, pz, ri, maxd, gap, L, inv_l, vx, vy, vz);
          if (m >= lo && m < hi) {
            s_v[grp][m - lo][0] = vx;
            s_v[grp][m - lo][1] = vy;
            s_v[grp][m - lo][2] = vz;
          }
        }
      } else {
#pragma unroll
        for (int u = 0; u < CPL; ++u)
          for (int k = st[u]; k < st[u] + ct[u]; ++k) {
            const float4 q = pr[k];
            if (k != kk && pbd_contact<U>(q.x, q.y, q.z, q.w, px, py, pz, ri, gap, L, inv_l) &&
                pbd_hit<U>(q.x, q.y, q.z, q.w, px, py, pz, ri, maxd, gap, L, inv_l, vx, vy,
                           vz)) {
              if (m >= lo && m < hi) {
                s_v[grp][m - lo][0] = vx;
                s_v[grp][m - lo][1] = vy;
                s_v[grp][m - lo][2] = vz;
              }
              ++m;
            }
          }
      }
    }
    __syncthreads();
    if (live && l == 0) {
      const int e = H - lo < HITS ? H - lo : HITS;
      for (int q = 0; q < e; ++q) {
        ax = ax + s_v[grp][q][0];
        ay = ay + s_v[grp][q][1];
        az = az + s_v[grp][q][2];
      }
    }
    if (r + 1 < rounds) __syncthreads();
  }
  if (!live || l != 0) return;
  const float maxt = maxd * dmul;
  const float cm = sqrtf(dot3(ax, ay, az));
  if (cm > maxt) {
    const float sc = maxt / cm;
    ax = ax * sc;
    ay = ay * sc;
    az = az * sc;
  }
  float w0 = gwrap(px + ax, L, inv_l), w1 = gwrap(py + ay, L, inv_l),
        w2 = gwrap(pz + az, L, inv_l);
  if (!BATCH && cell_of) {
    // the cell step of the rebuild that follows (see k_pbd_wave)
    w0 = gwrap(w0, L, inv_l);
    w1 = gwrap(w1, L, inv_l);
    w2 = gwrap(w2, L, inv_l);
    cell_of[i] = hash_axis(w0, G, half_l, inv_l) * G * G + hash_axis(w1, G, half_l, inv_l) * G +
                 hash_axis(w2, G, half_l, inv_l);
  }
  pout[3 * i] = w0;
  pout[3 * i + 1] = w1;
  pout[3 * i + 2] = w2;
}

// Lanes per particle of the slot-record PBD pass (a speed parameter; 0:
// 32 lanes (one stencil cell each) for single foams, 8 (four cells each)
// for batches of more than 131,072 particles, where it measured 3% faster;
// a single clustered foam of 200,000 is 1.6% faster with 32).
static int g_pbd_lanes = 0;

// Launches k_pbd_pr with g_pbd_lanes lanes per particle.
template <bool BATCH>
static void launch_pbd_pr(hipStream_t st, int n, const float *pin, float *pout, const float *pr,
                          const float *rad, int G, float L, float gap, float mdf, float dmul,
                          const int32_t *cs, const int32_t *cc, const int32_t *ci,
                          const int32_t *pinst, const fos_inst *inst, const int32_t *npass,
                          int pass, int32_t *cell_of) {
  const float half_l = (float)(0.5 * (double)L), inv_l = (float)(1.0 / (double)L);
  const int lpp = g_pbd_lanes ? g_pbd_lanes : (BATCH && n >= 131072 ? 8 : 32), npb = 256 / lpp;
  const int per = ((n + npb - 1) / npb + FOS_XCDS - 1) / FOS_XCDS;
#define PBD_PR(LP, U)                                                                            \
  hipLaunchKernelGGL((k_pbd_pr<BATCH, LP, U>), dim3((unsigned)(per * FOS_XCDS)), dim3(256), 0,  \
                     st, pin, pout, (const float4 *)pr, rad, n, G, L, half_l, inv_l, gap, mdf,    \
                     dmul, cs, cc, ci, pinst, inst, npass, pass, per, cell_of)
  const bool u = L == 1.0f;
  if (lpp == 8) {
    if (u) PBD_PR(8, true); else PBD_PR(8, false);
  } else if (lpp == 16) {
    if (u) PBD_PR(16, true); else PBD_PR(16, false);
  } else {
    if (u) PBD_PR(32, true); else PBD_PR(32, false);
  }
#undef PBD_PR
}

namespace fos {
int pbd_lanes_set(int lanes) {
  FOS_REQUIRE(lanes == 0 || lanes == 8 || lanes == 16 || lanes == 32,
              "pbd_lanes must be 0, 8, 16 or 32");
  g_pbd_lanes = lanes;
  return 0;
}
int pbd_lanes_get() { return g_pbd_lanes; }

// One PBD pass over the slot records of the rebuild before it, without the
// next rebuild's cell step (the pass before a step that moves the particles
// again, such as the OU jitter): what fos_pbd_pass computes, bit for bit.
int pbd_pass_records(const float *pos_in, float *pos_out, const float *rad, int n, int G,
                     float L, float gap_factor, float max_disp_frac, float disp_mult,
                     const int32_t *cell_start, const int32_t *cell_count,
                     const int32_t *cell_indices, const float *cell_pr, hipStream_t st) {
  FOS_REQUIRE(n >= 0 && G >= 3 && L > 0.0f &&
                  (n == 0 || (pos_in && pos_out && pos_in != pos_out && rad && cell_start &&
                              cell_count && cell_indices && cell_pr)),
              "pbd_pass_records: bad arguments (n=%d G=%d)", n, G);
  if (n == 0) return 0;
  launch_pbd_pr<false>(st, n, pos_in, pos_out, cell_pr, rad, G, L, gap_factor, max_disp_frac,
                       disp_mult, cell_start, cell_count, cell_indices, nullptr, nullptr,
                       nullptr, 0, nullptr);
  FOS_CHECK_LAUNCH("k_pbd_pr");
  return 0;
}
}  // namespace fos

__global__ void __launch_bounds__(256)
    k_max_overlap_pr_b(const float *__restrict__ pos, const float4 *__restrict__ pr,
                       const float *__restrict__ rad, int n, float L, float half_l, float inv_l,
                       const int32_t *__restrict__ cs, const int32_t *__restrict__ cc,
                       const int32_t *__restrict__ ci, float *__restrict__ lmd, float *gmax,
                       const int32_t *__restrict__ pinst, const fos_inst *__restrict__ inst,
                       int per_xcd) {
  const int grp = (int)(threadIdx.x >> 5), l = (int)(threadIdx.x & 31);
  // particles in cell order (see k_pbd_pr)
  const int k8 = xcd_tile((int)blockIdx.x, per_xcd) * 8 + grp;
  const int i = k8 < n ? ci[k8] : n;
  float lm = 0.0f;
  int b = 0;
  if (i < n && l < 27) {
    b = pinst[i];
    const int G = inst[b].G;
    const int32_t *csb = cs + inst[b].c0, *ccb = cc + inst[b].c0;
    const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
    const float ri = rad[i];
    const int cx = centre_axis(px, G, L, half_l, inv_l);
    const int cy = centre_axis(py, G, L, half_l, inv_l);
    const int cz = centre_axis(pz, G, L, half_l, inv_l);
    const int id = wrap_fast(cx + l / 9 - 1, G) * G * G + wrap_fast(cy + (l / 3) % 3 - 1, G) * G +
                   wrap_fast(cz + l % 3 - 1, G);
    const int st = csb[id], e = st + ccb[id];
    for (int k = st; k < e; ++k) {
      const int j = ci[k];
      const float4 q = pr[k];
      if (!(i < j)) continue;
      const float dx = gwrap(q.x - px, L, inv_l);
      const float dy = gwrap(q.y - py, L, inv_l);
      const float dz = gwrap(q.z - pz, L, inv_l);
      const float dist = sqrtf(dot3(dx, dy, dz) + 1e-8f);
      const float td = ri + q.w;
      if (dist < td) lm = fmaxf(lm, td - dist);
    }
  }
  for (int o = 16; o > 0; o >>= 1) lm = fmaxf(lm, __shfl_xor(lm, o, 32));
  // the workgroup's eight particles merge per foam (usually one) before the
  // atomics: a foam's particles are contiguous, so one slot would otherwise
  // take one atomic per particle
  __shared__ float s_lm[8];
  __shared__ int s_b[8];
  if (l == 0) {
    s_lm[grp] = i < n ? lm : 0.0f;
    s_b[grp] = i < n ? pinst[i] : -1;
    if (i < n) lmd[i] = lm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int cb = s_b[0];
    float cm = s_lm[0];
    for (int g = 1; g <= 8; ++g) {
      if (g == 8 || s_b[g] != cb) {
        if (cb >= 0 && cm > 0.0f) atomicMax((int *)&gmax[cb], __float_as_int(cm));
        if (g == 8) break;
        cb = s_b[g];
        cm = s_lm[g];
      } else {
        cm = fmaxf(cm, s_lm[g]);
      }
    }
  }
}

// Per-foam frame control on the device (a captured frame has no host
// round trip), with the host driver's double-precision arithmetic
// (run.py:709-751 pass count and global backpressure, run.py:1184-1231
// warm-start gate and settle counter).
__global__ void k_batch_control_pre(int B, const float *__restrict__ maxd,
                                    const double *__restrict__ r_max, int pbd_base, int pbd_max,
                                    double pbd_scale, double contact_tol,
                                    int32_t *__restrict__ settle, int32_t *__restrict__ passes,
                                    int32_t *__restrict__ active, float *__restrict__ back) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const double md = (double)maxd[b];
  const int p = (int)((double)pbd_base + pbd_scale * (md / (0.2 * r_max[b])));
  passes[b] = max(pbd_base, min(pbd_max, p));
  back[b] = (float)fmax(0.25, fmin(1.0, (contact_tol - md) / contact_tol));
  const int s = settle[b];
  active[b] = s > 0 ? 1 : 0;
  if (s > 0) settle[b] = s - 1;
}

__global__ void k_batch_control_post(const fos_inst *__restrict__ inst, int B,
                                     const int32_t *__restrict__ zeros,
                                     const int32_t *__restrict__ frame_dev, int warmstart,
                                     double zero_rate, int adjustment, int ran,
                                     int32_t *__restrict__ settle, int32_t *__restrict__ growth) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  int g = 1;
  if (ran) {
    const double f0 = (double)zeros[b] / (double)max(1, inst[b].n);
    g = (*frame_dev < warmstart || f0 > zero_rate) ? 0 : 1;
  }
  growth[b] = g;
  if (g) settle[b] = adjustment;
}

__global__ void k_frame_advance(int32_t *__restrict__ frame_dev) { *frame_dev += 1; }

extern "C" {

int fos_batch_max_overlap(fos_batch *bt, const float *pos, const float *rad, float L,
                          const int32_t *cell_start, const int32_t *cell_count,
                          const int32_t *cell_indices, const float *cell_pr,
                          float *local_max_depth, float *max_depth, void *stream) {
  FOS_REQUIRE(bt && pos && rad && cell_start && cell_count && cell_indices &&
                  local_max_depth && max_depth && L > 0.0f,
              "fos_batch_max_overlap: bad arguments");
  int rc = fos_fill_f32(max_depth, bt->B, 0.0f, stream);
  if (rc) return rc;
  const int n = bt->n_total;
  const float half_l = (float)(0.5 * (double)L), inv_l = (float)(1.0 / (double)L);
  if (cell_pr)
    hipLaunchKernelGGL(k_max_overlap_pr_b,
                       dim3((unsigned)((((n + 7) / 8 + FOS_XCDS - 1) / FOS_XCDS) * FOS_XCDS)),
                       dim3(256), 0, (hipStream_t)stream, pos, (const float4 *)cell_pr, rad, n,
                       L, half_l, inv_l, cell_start, cell_count, cell_indices, local_max_depth,
                       max_depth, bt->pinst, bt->inst, ((n + 7) / 8 + FOS_XCDS - 1) / FOS_XCDS);
  else
    hipLaunchKernelGGL(k_max_overlap_wave<true>, dim3((unsigned)((n + 7) / 8)), dim3(256), 0,
                       (hipStream_t)stream, pos, rad, n, 3, L, half_l, inv_l, cell_start,
                       cell_count, cell_indices, local_max_depth, max_depth, bt->pinst,
                       bt->inst);
  FOS_CHECK_LAUNCH("k_max_overlap_b");
  return 0;
}

int fos_batch_pbd_pass(fos_batch *bt, const float *pos_in, float *pos_out, const float *rad,
                       float L, float gap_factor, float max_disp_frac, float disp_mult,
                       const int32_t *cell_start, const int32_t *cell_count,
                       const int32_t *cell_indices, const float *cell_pr,
                       const int32_t *passes, int pass, void *stream) {
  FOS_REQUIRE(bt && pos_in && pos_out && pos_in != pos_out && rad && cell_start && cell_count &&
                  cell_indices && passes && L > 0.0f,
              "fos_batch_pbd_pass: bad arguments");
  const int n = bt->n_total;
  const float half_l = (float)(0.5 * (double)L), inv_l = (float)(1.0 / (double)L);
  if (cell_pr)
    launch_pbd_pr<true>((hipStream_t)stream, n, pos_in, pos_out, cell_pr, rad, 3, L, gap_factor,
                        max_disp_frac, disp_mult, cell_start, cell_count, cell_indices, bt->pinst,
                        bt->inst, passes, pass, nullptr);
  else
    hipLaunchKernelGGL(k_pbd_wave<true>, dim3((unsigned)((n + 7) / 8)), dim3(256), 0,
                       (hipStream_t)stream, pos_in, pos_out, rad, n, 3, L, half_l, inv_l,
                       gap_factor, max_disp_frac, disp_mult, cell_start, cell_count,
                       cell_indices, bt->pinst, bt->inst, passes, pass, (int32_t *)nullptr);
  FOS_CHECK_LAUNCH("k_pbd_b");
  return 0;
}

int fos_batch_nudge_radii(fos_batch *bt, float *rad, const float *rad_target,
                          const float *local_max_depth, float alpha, float max_step_pct,
                          int mode, float contact_tol, const float *back,
                          const int32_t *active, void *stream) {
  FOS_REQUIRE(bt && rad && rad_target && local_max_depth && back && active,
              "fos_batch_nudge_radii: bad arguments");
  const int n = bt->n_total;
  hipLaunchKernelGGL(k_nudge_b, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, (hipStream_t)stream,
                     rad, rad_target, local_max_depth, bt->pinst, bt->inst, active, back, n,
                     alpha, max_step_pct, mode, contact_tol);
  FOS_CHECK_LAUNCH("k_nudge_b");
  return 0;
}

int fos_batch_fsc_targets(fos_batch *bt, const float *rad, float *rad_target,
                          const int32_t *fsc, const float *fsc_ema, float growth_pct,
                          float deadband, const int32_t *active, void *stream) {
  FOS_REQUIRE(bt && rad && rad_target && fsc && fsc_ema && active,
              "fos_batch_fsc_targets: bad arguments");
  const int n = bt->n_total;
  hipLaunchKernelGGL(k_fsc_targets_b, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream, rad, rad_target, fsc, fsc_ema, bt->pinst, bt->inst,
                     active, n, growth_pct, deadband);
  FOS_CHECK_LAUNCH("k_fsc_targets_b");
  return 0;
}

int fos_batch_pressure_run(fos_batch *bt, fos_ctx *ctx, int frame, const int32_t *frame_dev,
                           int iters, float *rad, float *delta_r, const int32_t *face_ids,
                           const int32_t *fsc, float k, float p_exp, float pair_cap, int max_nei,
                           void *stream) {
  FOS_REQUIRE(bt && ctx && bt->n_total <= ctx->max_n && iters >= 0 && rad && delta_r &&
                  face_ids && fsc,
              "fos_batch_pressure_run: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const int n = bt->n_total;
  if (iters == 0) return 0;
  float *a = rad, *b = ctx->rad_tmp;
  for (int it = 0; it < iters; ++it) {
    hipLaunchKernelGGL(k_pressure_step_b, dim3((unsigned)((n + 15) / 16)), dim3(256), 0, st, n,
                       frame, frame_dev, a, b, delta_r, face_ids, fsc, bt->pinst, bt->inst, k,
                       p_exp, 1.0f / p_exp, pair_cap, max_nei);
    FOS_CHECK_LAUNCH("k_pressure_step_b");
    float *t = a;
    a = b;
    b = t;
  }
  if (a != rad) {
    hipLaunchKernelGGL(k_copy_f32, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, st, a, rad,
                       (int64_t)n);
    FOS_CHECK_LAUNCH("k_copy_f32");
  }
  return 0;
}

int fos_batch_brownian(fos_batch *bt, float *pos, float strength, float L, int frame,
                       const int32_t *frame_dev, int mode, void *stream) {
  FOS_REQUIRE(bt && pos && L > 0.0f && (mode == 0 || mode == 1),
              "fos_batch_brownian: bad arguments");
  const int n = bt->n_total;
  hipLaunchKernelGGL(k_brownian_b, dim3(blocks_for(n)), dim3(FOS_BLOCK), 0, (hipStream_t)stream,
                     pos, bt->pinst, bt->inst, n, strength, L, L * 0.5f, frame, frame_dev, mode);
  FOS_CHECK_LAUNCH("k_brownian_b");
  return 0;
}

int fos_batch_count_zero(fos_batch *bt, const int32_t *fsc, int32_t *out, void *stream) {
  FOS_REQUIRE(bt && fsc && out, "fos_batch_count_zero: bad arguments");
  int rc = fos_fill_i32(out, bt->B, 0, stream);
  if (rc) return rc;
  hipLaunchKernelGGL(k_count_zero_b, dim3(blocks_for(bt->n_total)), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream, fsc, bt->pinst, bt->n_total, out);
  FOS_CHECK_LAUNCH("k_count_zero_b");
  return 0;
}

int fos_batch_control_pre(fos_batch *bt, const float *max_depth, const double *r_max,
                          int pbd_base, int pbd_max, double pbd_scale, double contact_tol,
                          int32_t *settle, int32_t *passes, int32_t *active, float *back,
                          void *stream) {
  FOS_REQUIRE(bt && max_depth && r_max && settle && passes && active && back && contact_tol != 0.0,
              "fos_batch_control_pre: bad arguments");
  hipLaunchKernelGGL(k_batch_control_pre, dim3(blocks_for(bt->B)), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream, bt->B, max_depth, r_max, pbd_base, pbd_max, pbd_scale,
                     contact_tol, settle, passes, active, back);
  FOS_CHECK_LAUNCH("k_batch_control_pre");
  return 0;
}

int fos_batch_control_post(fos_batch *bt, const int32_t *zeros, const int32_t *frame_dev,
                           int warmstart_frames, double fsc_zero_rate, int adjustment_frames,
                           int ran, int32_t *settle, int32_t *growth, void *stream) {
  FOS_REQUIRE(bt && zeros && frame_dev && settle && growth,
              "fos_batch_control_post: bad arguments");
  hipLaunchKernelGGL(k_batch_control_post, dim3(blocks_for(bt->B)), dim3(FOS_BLOCK), 0,
                     (hipStream_t)stream, bt->inst, bt->B, zeros, frame_dev, warmstart_frames,
                     fsc_zero_rate, adjustment_frames, ran, settle, growth);
  FOS_CHECK_LAUNCH("k_batch_control_post");
  return 0;
}

int fos_batch_frame_advance(int32_t *frame_dev, void *stream) {
  FOS_REQUIRE(frame_dev, "fos_batch_frame_advance: bad arguments");
  hipLaunchKernelGGL(k_frame_advance, dim3(1), dim3(1), 0, (hipStream_t)stream, frame_dev);
  FOS_CHECK_LAUNCH("k_frame_advance");
  return 0;
}

}  // extern "C"
