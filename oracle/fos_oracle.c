/*
 * fos_oracle.c -- CPU restatement of the reference per-frame foam step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the
 * timed CPU baseline.  Nothing in the product links or loads this file.
 *
 * Each function follows one reference Taichi kernel as if its outermost loops
 * ran on one thread in program order (the reference's serial semantics); the
 * cited file:line is where the loop lives.  Loops whose iterations are
 * independent (per voxel or per particle, writing only their own element)
 * carry an OpenMP pragma; that cannot change results.  Order-defined loops
 * stay serial.  Compile with -ffp-contract=off, without fast-math.
 */
#include "fos_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXN ORC_MAX_NEIGHBORS

void orc_set_threads(int nthreads) {
#ifdef _OPENMP
  omp_set_num_threads(nthreads > 0 ? nthreads : 1);
#else
  (void)nthreads;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Geometry helpers                                                          */
/* ------------------------------------------------------------------------ */

/* jfa.py:247-251 wrap_voxel_idx / grid.py:91-95 wrap_cell */
static inline int wrapi(int i, int r) { return ((i % r) + r) % r; }

/* jfa.py:303-308 wrapP: d - L*round(d/L) */
static inline float jwrap(float d, float L) { return d - L * roundf(d / L); }

/* jfa.py:285-287 voxel_to_world */
static inline float vcenter(int i, float voxel, float half_l) {
  return ((float)i + 0.5f) * voxel - half_l;
}

/* jfa.py:451-477 power_distance_sq: dot(d,d) - (beta*r)^2, dot summed
 * left to right ((x*x + y*y) + z*z). */
float orc_power_distance(float cx, float cy, float cz, const float *p, float r,
                         float beta, float L) {
  float dx = jwrap(cx - p[0], L);
  float dy = jwrap(cy - p[1], L);
  float dz = jwrap(cz - p[2], L);
  float d2 = dx * dx + dy * dy;
  d2 = d2 + dz * dz;
  float b = beta * r;
  return d2 - b * b;
}

/* jfa.py:255-270 world_to_voxel, linearised x*R^2 + y*R + z */
int orc_seed_voxel(const float *p, int res, float L) {
  float half_l = (float)(0.5 * (double)L);
  int c[3];
  for (int d = 0; d < 3; ++d) {
    float nrm = (p[d] + half_l) / L;
    c[d] = wrapi((int)floorf(nrm * (float)res), res);
  }
  return (c[0] * res + c[1]) * res + c[2];
}

/* grid.py:30-48 wrapP: p - L*floor(p*INV_L + 0.5) */
static inline float gwrap(float p, float L, float inv_l) {
  return p - L * floorf(p * inv_l + 0.5f);
}

/* ------------------------------------------------------------------------ */
/* JFA                                                                       */
/* ------------------------------------------------------------------------ */

void orc_jfa_init(const float *pos, const float *rad, int n, int res, float L,
                  float voxel, float beta, int32_t *grid, int32_t *temp) {
  const int64_t V = (int64_t)res * res * res;
  const float half_l = (float)(0.5 * (double)L);
  for (int64_t v = 0; v < V; ++v) {
    grid[v] = -1;
    temp[v] = -1;
  }
  /* jfa.py:503-520, serial: first seeder takes the voxel, later ones replace
   * it only with a strictly smaller power distance. */
  for (int i = 0; i < n; ++i) {
    int v = orc_seed_voxel(pos + 3 * i, res, L);
    int cur = grid[v];
    if (cur == -1) {
      grid[v] = i;
    } else {
      int z = v % res, y = (v / res) % res, x = v / (res * res);
      float cx = vcenter(x, voxel, half_l), cy = vcenter(y, voxel, half_l),
            cz = vcenter(z, voxel, half_l);
      float pc = orc_power_distance(cx, cy, cz, pos + 3 * cur, rad[cur], beta, L);
      float pn = orc_power_distance(cx, cy, cz, pos + 3 * i, rad[i], beta, L);
      if (pn < pc) grid[v] = i;
    }
  }
}

int64_t orc_jfa_step(const float *pos, const float *rad, int res, float L,
                     float voxel, float beta, int step, const int32_t *in,
                     int32_t *out) {
  const float half_l = (float)(0.5 * (double)L);
  int64_t changes = 0;
#pragma omp parallel for collapse(2) reduction(+ : changes) schedule(static)
  for (int x = 0; x < res; ++x) {
    for (int y = 0; y < res; ++y) {
      float cx = vcenter(x, voxel, half_l), cy = vcenter(y, voxel, half_l);
      for (int z = 0; z < res; ++z) {
        float cz = vcenter(z, voxel, half_l);
        int64_t v = ((int64_t)x * res + y) * res + z;
        int old = in[v];
        int best = old;
        float bd = 1e10f;
        if (best >= 0)
          bd = orc_power_distance(cx, cy, cz, pos + 3 * best, rad[best], beta, L);
        /* jfa.py:552-565: dx outermost, strict < keeps the earliest winner */
        for (int dx = -1; dx <= 1; ++dx)
          for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
              int nx = wrapi(x + dx * step, res), ny = wrapi(y + dy * step, res),
                  nz = wrapi(z + dz * step, res);
              int c = in[((int64_t)nx * res + ny) * res + nz];
              if (c >= 0) {
                float pd = orc_power_distance(cx, cy, cz, pos + 3 * c, rad[c], beta, L);
                if (pd < bd) {
                  bd = pd;
                  best = c;
                }
              }
            }
        out[v] = best;
        if (best != old) changes += 1;
      }
    }
  }
  return changes;
}

int orc_jfa_run(const float *pos, const float *rad, int n, int res, float L,
                float voxel, float beta, int num_passes, int early_exit_enabled,
                double th_pass_delta, int min_passes, int32_t *grid,
                int32_t *temp, int64_t *changes_hist, int *early_exit) {
  (void)n;
  const int64_t V = (int64_t)res * res * res;
  int step = res / 2;
  int ping = 0, passes = 0;
  *early_exit = 0;
  for (int pass = 0; pass < num_passes; ++pass) {
    if (step < 1) step = 1;
    int64_t ch;
    if (ping == 0)
      ch = orc_jfa_step(pos, rad, res, L, voxel, beta, step, grid, temp);
    else
      ch = orc_jfa_step(pos, rad, res, L, voxel, beta, step, temp, grid);
    ping = 1 - ping;
    step = step / 2;
    passes += 1;
    if (changes_hist) changes_hist[pass] = ch;
    /* jfa.py:1104-1112: change fraction in double, strict < */
    double frac = (double)ch / (double)(V > 1 ? V : 1);
    if (early_exit_enabled && pass >= (min_passes - 1) && frac < th_pass_delta) {
      *early_exit = 1;
      break;
    }
  }
  if (ping == 1) memcpy(grid, temp, (size_t)V * sizeof(int32_t));
  return passes;
}

void orc_majority(const int32_t *grid, int32_t *filtered, int res) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int x = 0; x < res; ++x)
    for (int y = 0; y < res; ++y)
      for (int z = 0; z < res; ++z) {
        int64_t v = ((int64_t)x * res + y) * res + z;
        int my = grid[v];
        if (my < 0) {
          filtered[v] = -1;
          continue;
        }
        int ids[7], cnt[7], nu = 1;
        ids[0] = my;
        cnt[0] = 1;
        const int c[3] = {x, y, z};
        for (int axis = 0; axis < 3; ++axis)
          for (int dir = -1; dir <= 1; dir += 2) {
            int q[3] = {c[0], c[1], c[2]};
            q[axis] = wrapi(q[axis] + dir, res);
            int nid = grid[((int64_t)q[0] * res + q[1]) * res + q[2]];
            if (nid >= 0) {
              int found = 0;
              for (int k = 0; k < nu; ++k)
                if (ids[k] == nid) {
                  cnt[k] += 1;
                  found = 1;
                  break;
                }
              if (!found && nu < 7) {
                ids[nu] = nid;
                cnt[nu] = 1;
                nu += 1;
              }
            }
          }
        int maxc = 0, best = my;
        for (int k = 0; k < nu; ++k)
          if (cnt[k] > maxc) {
            maxc = cnt[k];
            best = ids[k];
          }
        filtered[v] = best;
      }
}

void orc_accumulate_faces(const int32_t *filtered, int res, int n,
                          int32_t *face_ids, int32_t *face_counts) {
  for (int64_t i = 0; i < (int64_t)n * MAXN; ++i) {
    face_ids[i] = -1;
    face_counts[i] = 0;
  }
  /* jfa.py:723-762 in serial program order: x, y, z row-major, then axis
   * 0..2, direction -1 then +1.  A pair (a<b) lives in row a; the first free
   * slot is claimed on first encounter; a full row drops new partners. */
  for (int x = 0; x < res; ++x)
    for (int y = 0; y < res; ++y)
      for (int z = 0; z < res; ++z) {
        int a = filtered[((int64_t)x * res + y) * res + z];
        if (a < 0 || a >= n) continue;
        const int c[3] = {x, y, z};
        for (int axis = 0; axis < 3; ++axis)
          for (int dir = -1; dir <= 1; dir += 2) {
            int q[3] = {c[0], c[1], c[2]};
            q[axis] = wrapi(q[axis] + dir, res);
            int b = filtered[((int64_t)q[0] * res + q[1]) * res + q[2]];
            if (b >= 0 && b < n && a < b) {
              int32_t *ids = face_ids + (int64_t)a * MAXN;
              int32_t *cts = face_counts + (int64_t)a * MAXN;
              for (int s = 0; s < MAXN; ++s) {
                if (ids[s] == b) {
                  cts[s] += 1;
                  break;
                } else if (ids[s] == -1) {
                  ids[s] = b;
                  cts[s] += 1;
                  break;
                }
              }
            }
          }
      }
}

/* jfa.py:379-397 expected_area: 1.5 * ((2*(0.5*(ri+rj)))/voxel)^2 */
static inline float expected_area(float ri, float rj, float voxel) {
  float r_mean = 0.5f * (ri + rj);
  float t = (2.0f * r_mean) / voxel;
  return 1.5f * (t * t);
}

/* jfa.py:400-443 add_neighbor.  The "already present" branch cannot fire
 * under the serial order (row i holds lower partners then its own higher
 * partners, each once), but it is restated anyway. */
static void add_neighbor(int i, int j, float score, float *prev, int32_t *nl,
                         int32_t *nc, int32_t *used, int32_t *fsc) {
  int u = used[i];
  for (int s = 0; s < u; ++s)
    if (nl[(int64_t)i * MAXN + s] == j) {
      prev[(int64_t)i * MAXN + s] = score;
      return;
    }
  if (u < MAXN) {
    nl[(int64_t)i * MAXN + u] = j;
    prev[(int64_t)i * MAXN + u] = score;
    nc[i] += 1;
    fsc[i] += 1;
    used[i] += 1;
  }
}

void orc_finalize_faces(int n, const float *rad, float voxel,
                        const int32_t *face_ids, const int32_t *face_counts,
                        float *prev_scores, int32_t *neighbor_list,
                        int32_t *neighbor_count, int32_t *used_count,
                        int32_t *fsc) {
  for (int i = 0; i < n; ++i) {
    fsc[i] = 0;
    neighbor_count[i] = 0;
    used_count[i] = 0;
    for (int s = 0; s < MAXN; ++s) neighbor_list[(int64_t)i * MAXN + s] = -1;
  }
  const float T_ADD = 0.15f, T_DROP = 0.10f; /* jfa.py:64-65 */
  for (int i = 0; i < n; ++i) {
    float ri = rad[i];
    for (int s = 0; s < MAXN; ++s) {
      int j = face_ids[(int64_t)i * MAXN + s];
      if (j < 0 || j >= n) continue;
      int v = face_counts[(int64_t)i * MAXN + s];
      if (v <= 0) continue;
      float e = expected_area(ri, rad[j], voxel);
      if (e <= 0.0f) continue;
      float score = (float)v / e;
      float ps = prev_scores[(int64_t)i * MAXN + s];
      int keep = (ps >= T_DROP && score >= T_DROP);
      int add = (score >= T_ADD);
      if (add || keep) {
        add_neighbor(i, j, score, prev_scores, neighbor_list, neighbor_count,
                     used_count, fsc);
        add_neighbor(j, i, score, prev_scores, neighbor_list, neighbor_count,
                     used_count, fsc);
      } else {
        prev_scores[(int64_t)i * MAXN + s] = 0.0f;
      }
    }
  }
}

void orc_fsc_ema(int n, float alpha, const int32_t *fsc, float *fsc_ema) {
  for (int i = 0; i < n; ++i) {
    if (fsc_ema[i] == 0.0f && fsc[i] > 0)
      fsc_ema[i] = (float)fsc[i];
    else
      fsc_ema[i] = alpha * (float)fsc[i] + (1.0f - alpha) * fsc_ema[i];
  }
}

void orc_jfa_stats(int n, const float *rad, float voxel, const int32_t *fsc,
                   const float *fsc_ema, const int32_t *used_count,
                   const int32_t *face_ids, const int32_t *face_counts,
                   int64_t *out_i, double *out_f) {
  int64_t p_dir = 0, overflow = 0, sum_fsc = 0;
  int64_t mn = 999999, mx = -1;
  double sum_ema = 0.0;
  for (int i = 0; i < n; ++i) {
    p_dir += used_count[i];                   /* jfa.py:1215-1220 */
    if (used_count[i] >= MAXN) overflow += 1; /* jfa.py:1231-1235 */
    sum_fsc += fsc[i];                        /* jfa.py:1251-1259 */
    sum_ema += fsc_ema[i];
    if (fsc[i] < mn) mn = fsc[i];
    if (fsc[i] > mx) mx = fsc[i];
  }
  /* jfa.py:1286-1328: iterates range(used) over face_ids (reference quirk) */
  double ssum = 0.0;
  int64_t scount = 0;
  for (int pass = 0; pass < 2; ++pass) {
    double mean = scount ? ssum / (double)scount : 0.0;
    double sq = 0.0;
    for (int i = 0; i < n; ++i) {
      int u = used_count[i];
      for (int s = 0; s < u; ++s) {
        int j = face_ids[(int64_t)i * MAXN + s];
        if (j < 0 || j >= n) continue;
        int v = face_counts[(int64_t)i * MAXN + s];
        if (v <= 0) continue;
        float e = expected_area(rad[i], rad[j], voxel);
        if (e > 0.0f) {
          float sc = (float)v / e;
          if (pass == 0) {
            ssum += sc;
            scount += 1;
          } else {
            sq += ((double)sc - mean) * ((double)sc - mean);
          }
        }
      }
    }
    if (pass == 1) {
      out_f[2] = mean;
      out_f[3] = scount ? sqrt(sq / (double)scount) : 0.0;
    }
  }
  out_i[0] = p_dir;
  out_i[1] = overflow;
  out_i[2] = mn;
  out_i[3] = mx;
  out_i[4] = sum_fsc;
  out_f[0] = n ? (double)sum_fsc / (double)n : 0.0;
  out_f[1] = n ? sum_ema / (double)n : 0.0;
}

void orc_validate(int n, const int32_t *nl, const int32_t *nc, int64_t *out) {
  int64_t asym = 0, over = 0, self = 0;
  for (int i = 0; i < n; ++i) {
    for (int s = 0; s < MAXN; ++s) {
      int j = nl[(int64_t)i * MAXN + s];
      if (j == i) self += 1; /* jfa.py:1392-1395 */
      if (j >= 0 && j < n) { /* jfa.py:1347-1360 */
        int found = 0;
        for (int k = 0; k < MAXN; ++k)
          if (nl[(int64_t)j * MAXN + k] == i) {
            found = 1;
            break;
          }
        if (!found) asym += 1;
      }
    }
    if (nc[i] >= MAXN) over += 1; /* jfa.py:1375-1377 */
  }
  out[0] = asym;
  out[1] = over;
  out[2] = self;
}

/* jfa.py:316-352 world_to_tile + wrap_tile_idx; jfa.py:983-1008 */
void orc_mark_dirty_tiles(const float *pos, const float *rad,
                          const float *pos_prev, const float *rad_prev, int n,
                          int res, float L, float voxel, int tile, int halo,
                          float pos_thr, float rad_thr, uint8_t *tile_dirty) {
  const int tpa = (res + tile - 1) / tile;
  for (int i = 0; i < n; ++i) {
    float dx = pos[3 * i] - pos_prev[3 * i], dy = pos[3 * i + 1] - pos_prev[3 * i + 1],
          dz = pos[3 * i + 2] - pos_prev[3 * i + 2];
    float pd = sqrtf((dx * dx + dy * dy) + dz * dz);
    float rd = fabsf(rad[i] - rad_prev[i]);
    int vc = orc_seed_voxel(pos + 3 * i, res, L);
    int vp = orc_seed_voxel(pos_prev + 3 * i, res, L);
    int tc[3] = {(vc / (res * res)) / tile, ((vc / res) % res) / tile, (vc % res) / tile};
    int tp[3] = {(vp / (res * res)) / tile, ((vp / res) % res) / tile, (vp % res) / tile};
    int dirty = (pd > pos_thr * voxel) || (rd > rad_thr * voxel) || tc[0] != tp[0] ||
                tc[1] != tp[1] || tc[2] != tp[2];
    if (!dirty) continue;
    for (int a = -halo; a <= halo; ++a)
      for (int b = -halo; b <= halo; ++b)
        for (int c = -halo; c <= halo; ++c) {
          int t0 = wrapi(tc[0] + c, tpa), t1 = wrapi(tc[1] + b, tpa),
              t2 = wrapi(tc[2] + a, tpa);
          tile_dirty[(t0 * tpa + t1) * tpa + t2] = 1;
        }
  }
}

/* ------------------------------------------------------------------------ */
/* Spatial hash grid                                                         */
/* ------------------------------------------------------------------------ */

/* grid.py:202-219: cell of a wrapped position, with the 1e-7 hash epsilon */
static inline int hash_cell(const float *pw, int G, float half_l, float inv_l) {
  int c[3];
  for (int d = 0; d < 3; ++d) {
    float q = ((pw[d] + half_l) + 1e-7f) * inv_l;
    c[d] = wrapi((int)floorf(q * (float)G), G);
  }
  return c[0] * G * G + c[1] * G + c[2];
}

void orc_grid_build(float *pos, int n, int G, float L, int32_t *cell_count,
                    int32_t *cell_start, int32_t *cell_indices) {
  const int C = G * G * G;
  const float half_l = (float)(0.5 * (double)L), inv_l = (float)(1.0 / (double)L);
  for (int c = 0; c < C; ++c) cell_count[c] = 0;
  for (int i = 0; i < n; ++i) { /* grid.py:202-222 */
    float pw[3] = {gwrap(pos[3 * i], L, inv_l), gwrap(pos[3 * i + 1], L, inv_l),
                   gwrap(pos[3 * i + 2], L, inv_l)};
    cell_count[hash_cell(pw, G, half_l, inv_l)] += 1;
  }
  cell_start[0] = 0; /* grid.py:241-243, serial */
  for (int c = 1; c < C; ++c) cell_start[c] = cell_start[c - 1] + cell_count[c - 1];
  int32_t *wr = (int32_t *)malloc((size_t)C * sizeof(int32_t));
  memcpy(wr, cell_start, (size_t)C * sizeof(int32_t));
  for (int i = 0; i < n; ++i) { /* grid.py:285-309, serial => stable */
    float pw[3] = {gwrap(pos[3 * i], L, inv_l), gwrap(pos[3 * i + 1], L, inv_l),
                   gwrap(pos[3 * i + 2], L, inv_l)};
    pos[3 * i] = pw[0];
    pos[3 * i + 1] = pw[1];
    pos[3 * i + 2] = pw[2];
    int c = hash_cell(pw, G, half_l, inv_l);
    cell_indices[wr[c]++] = i;
  }
  free(wr);
}

/* dynamics.py:101-103: stencil centre, no epsilon, truncating cast */
static inline void centre_cell(const float *p, int G, float L, float half_l,
                               float inv_l, int c[3]) {
  for (int d = 0; d < 3; ++d) {
    float pw = gwrap(p[d], L, inv_l);
    float q = (pw + half_l) * inv_l;
    c[d] = (int)(q * (float)G);
  }
}

/* One particle's PBD correction (dynamics.py:96-164).  pj_src is where j's
 * position is read from (pass-start snapshot for Jacobi, live for GS). */
static void pbd_one(int i, const float *pi_p, const float *pj_src,
                    const float *rad, int G, float L, float gap,
                    float max_disp_frac, float disp_mult, const int32_t *cs,
                    const int32_t *cc, const int32_t *ci, float corr[3]) {
  const float half_l = (float)(0.5 * (double)L), inv_l = (float)(1.0 / (double)L);
  float cx = 0.0f, cy = 0.0f, cz = 0.0f;
  int mc[3];
  centre_cell(pi_p, G, L, half_l, inv_l, mc);
  const float ri = rad[i];
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        int nx = wrapi(mc[0] + dx, G), ny = wrapi(mc[1] + dy, G), nz = wrapi(mc[2] + dz, G);
        int id = nx * G * G + ny * G + nz;
        int st = cs[id], ct = cc[id];
        for (int k = st; k < st + ct; ++k) {
          int j = ci[k];
          if (j == i) continue;
          const float *pj = pj_src + 3 * j;
          float ddx = gwrap(pj[0] - pi_p[0], L, inv_l);
          float ddy = gwrap(pj[1] - pi_p[1], L, inv_l);
          float ddz = gwrap(pj[2] - pi_p[2], L, inv_l);
          float d2 = ddx * ddx + ddy * ddy;
          d2 = d2 + ddz * ddz;
          float dist = sqrtf(d2 + 1e-12f);
          float ts = (ri + rad[j]) * gap;
          if (dist < ts) {
            float overlap = ts - dist;
            float ux = ddx / dist, uy = ddy / dist, uz = ddz / dist;
            float mag = overlap * 0.5f;
            float maxd = max_disp_frac * ri;
            mag = fminf(mag, maxd);
            float vx = -ux * mag, vy = -uy * mag, vz = -uz * mag;
            vx = fmaxf(-maxd, fminf(maxd, vx));
            vy = fmaxf(-maxd, fminf(maxd, vy));
            vz = fmaxf(-maxd, fminf(maxd, vz));
            cx = cx + vx;
            cy = cy + vy;
            cz = cz + vz;
          }
        }
      }
  float maxt = max_disp_frac * ri * disp_mult;
  float cm2 = cx * cx + cy * cy;
  cm2 = cm2 + cz * cz;
  float cm = sqrtf(cm2);
  if (cm > maxt) {
    float s = maxt / cm;
    cx = cx * s;
    cy = cy * s;
    cz = cz * s;
  }
  corr[0] = cx;
  corr[1] = cy;
  corr[2] = cz;
}

void orc_pbd_jacobi(const float *pos_in, float *pos_out, const float *rad,
                    int n, int G, float L, float gap_factor, float max_disp_frac,
                    float disp_mult, const int32_t *cs, const int32_t *cc,
                    const int32_t *ci) {
  const float inv_l = (float)(1.0 / (double)L);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    float corr[3];
    pbd_one(i, pos_in + 3 * i, pos_in, rad, G, L, gap_factor, max_disp_frac,
            disp_mult, cs, cc, ci, corr);
    for (int d = 0; d < 3; ++d)
      pos_out[3 * i + d] = gwrap(pos_in[3 * i + d] + corr[d], L, inv_l);
  }
}

void orc_pbd_gauss_seidel(float *pos, const float *rad, int n, int G, float L,
                          float gap_factor, float max_disp_frac, float disp_mult,
                          const int32_t *cs, const int32_t *cc,
                          const int32_t *ci) {
  const float inv_l = (float)(1.0 / (double)L);
  for (int i = 0; i < n; ++i) {
    float corr[3];
    pbd_one(i, pos + 3 * i, pos, rad, G, L, gap_factor, max_disp_frac, disp_mult,
            cs, cc, ci, corr);
    for (int d = 0; d < 3; ++d) pos[3 * i + d] = gwrap(pos[3 * i + d] + corr[d], L, inv_l);
  }
}

float orc_max_overlap(const float *pos, const float *rad, int n, int G, float L,
                      const int32_t *cs, const int32_t *cc, const int32_t *ci,
                      float *lmd) {
  const float half_l = (float)(0.5 * (double)L), inv_l = (float)(1.0 / (double)L);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    const float *pi_p = pos + 3 * i;
    float ri = rad[i];
    int mc[3];
    centre_cell(pi_p, G, L, half_l, inv_l, mc);
    float lm = 0.0f;
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = -1; dz <= 1; ++dz) {
          int id = wrapi(mc[0] + dx, G) * G * G + wrapi(mc[1] + dy, G) * G + wrapi(mc[2] + dz, G);
          for (int k = cs[id]; k < cs[id] + cc[id]; ++k) {
            int j = ci[k];
            if (i < j) {
              const float *pj = pos + 3 * j;
              float ddx = gwrap(pj[0] - pi_p[0], L, inv_l);
              float ddy = gwrap(pj[1] - pi_p[1], L, inv_l);
              float ddz = gwrap(pj[2] - pi_p[2], L, inv_l);
              float d2 = ddx * ddx + ddy * ddy;
              d2 = d2 + ddz * ddz;
              float dist = sqrtf(d2 + 1e-8f);
              float td = ri + rad[j];
              if (dist < td) lm = fmaxf(lm, td - dist);
            }
          }
        }
    lmd[i] = lm;
  }
  float g = 0.0f; /* dynamics.py:361-364, as a true max */
  for (int i = 0; i < n; ++i) g = fmaxf(g, lmd[i]);
  return g;
}

/* ------------------------------------------------------------------------ */
/* Controller, pressure, noise                                               */
/* ------------------------------------------------------------------------ */

/* dynamics.py:1226-1240 */
static inline float smoothstep(float e0, float e1, float x) {
  float t = fminf(1.0f, fmaxf(0.0f, (x - e0) / (e1 - e0)));
  return t * t * (3.0f - 2.0f * t);
}

void orc_fsc_targets(const float *rad, float *rad_target, const int32_t *fsc,
                     const float *fsc_ema, int n, int f_low, int f_high,
                     float growth_pct, float r_min, float r_max,
                     float deadband) {
  const float db = deadband, gap = 1.0f, idle = 0.003f;
  const float fl = (float)f_low, fh = (float)f_high;
  for (int i = 0; i < n; ++i) {
    float r0 = rad[i];
    float f = fsc_ema[i] > 0.0f ? fsc_ema[i] : (float)fsc[i];
    float gain;
    if (f < fl - gap) {
      gain = 1.0f;
      rad_target[i] = fminf(r0 * (1.0f + growth_pct * gain), r_max);
    } else if (f < fl) {
      gain = 1.0f - smoothstep(fl - gap - db, fl, f);
      rad_target[i] = fminf(r0 * (1.0f + growth_pct * gain), r_max);
    } else if (f > fh + gap) {
      gain = 1.0f;
      rad_target[i] = fmaxf(r0 * (1.0f - growth_pct * gain), r_min);
    } else if (f > fh) {
      gain = smoothstep(fh, fh + gap + db, f);
      rad_target[i] = fmaxf(r0 * (1.0f - growth_pct * gain), r_min);
    } else {
      float center = 0.5f * (fl + fh);
      if (f < center)
        rad_target[i] = fminf(r0 * (1.0f + growth_pct * idle), r_max);
      else
        rad_target[i] = fmaxf(r0 * (1.0f - growth_pct * idle), r_min);
    }
  }
}

void orc_controller_stats(const float *rad, const float *rad_target,
                          const float *fsc_ema, int n, int f_low, int f_high,
                          int64_t *out_i, double *out_f) {
  int64_t g = 0, s = 0, idl = 0;
  double sum = 0.0;
  float mx = 0.0f;
  const float fl = (float)f_low, fh = (float)f_high;
  for (int i = 0; i < n; ++i) {
    float f = fsc_ema[i];
    float dr = rad_target[i] - rad[i];
    if (f < fl)
      g += 1;
    else if (f > fh)
      s += 1;
    else
      idl += 1;
    sum += dr;
    if (fabsf(dr) > mx) mx = fabsf(dr);
  }
  out_i[0] = g;
  out_i[1] = s;
  out_i[2] = idl;
  out_f[0] = sum / (double)(n > 1 ? n : 1);
  out_f[1] = mx;
}

void orc_nudge_radii(float *rad, const float *rad_target, const float *lmd,
                     int n, float alpha, float max_step_pct, float r_min,
                     float r_max, float back_global, int mode,
                     float contact_tol) {
  for (int i = 0; i < n; ++i) {
    float r0 = rad[i];
    float rt = fminf(r_max, fmaxf(r_min, rad_target[i]));
    float d = alpha * (rt - r0);
    float b = 1.0f;
    if (mode == 1) {
      b = back_global;
    } else if (mode == 2) {
      b = (contact_tol - lmd[i]) / contact_tol;
      b = fminf(1.0f, fmaxf(0.25f, b));
    }
    d = d * b;
    float cap = max_step_pct * r0;
    if (d > cap) d = cap;
    if (d < -cap) d = -cap;
    rad[i] = fminf(r_max, fmaxf(r_min, r0 + d));
  }
}

/* dynamics.py:1451: int32 arithmetic with two's-complement wraparound */
int32_t orc_hash(int32_t i, int32_t frame) {
  uint32_t t = (uint32_t)i + 12345u * (uint32_t)frame;
  uint32_t h = 1103515245u * t + 12345u;
  return (int32_t)(h & 0x7fffffffu);
}

/* fp32 power evaluated in double and rounded once (the correctly rounded
 * fp32 result in all but ~2^-29 of cases); the product uses the same form. */
static inline float pow_f32(float x, float y) { return (float)pow((double)x, (double)y); }

/* r^p for volumes: for p == 3, (r*r)*r in double is the correctly rounded
 * cube (r*r is exact), which is what a correctly rounded pow returns. */
static inline float vol_f32(float r, float p) {
  if (p == 3.0f) {
    double d = (double)r;
    return (float)(d * d * d);
  }
  return pow_f32(r, p);
}

float orc_pressure(int n, int frame, float *rad, float *delta_r,
                   const int32_t *face_ids, const int32_t *fsc, float k,
                   float p_exp, float pair_cap, int max_nei, float r_min,
                   float r_max, int32_t *changed) {
  const float inv_p = 1.0f / p_exp;
  for (int i = 0; i < n; ++i) delta_r[i] = 0.0f;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    float vi = vol_f32(rad[i], p_exp);
    int fc = fsc[i];
    if (fc <= 0) continue;
    int an = fc < max_nei ? fc : max_nei;
    int start = orc_hash(i, frame) % fc;
    float acc = 0.0f;
    for (int kk = 0; kk < an; ++kk) {
      int idx = (start + kk) % fc;
      int j = face_ids[(int64_t)i * MAXN + idx];
      if (j < 0 || j >= n || j == i) continue;
      float vj = vol_f32(rad[j], p_exp);
      float dv_raw = k * (vi - vj);
      float vmin = fminf(vi, vj);
      float dv = fmaxf(-pair_cap * vmin, fminf(pair_cap * vmin, dv_raw));
      float vnew = vi - dv;
      acc = acc + (pow_f32(vnew, inv_p) - rad[i]);
    }
    delta_r[i] = acc;
  }
  float mx = 0.0f;
  int32_t ch = 0;
  for (int i = 0; i < n; ++i) {
    float d = delta_r[i];
    if (fabsf(d) > 1e-12f) {
      ch += 1;
      if (fabsf(d) > mx) mx = fabsf(d);
    }
    rad[i] = fmaxf(r_min, fminf(r_max, rad[i] + d));
  }
  *changed = ch;
  return mx;
}

/* Philox4x32-10 (Salmon et al., "Parallel random numbers: as easy as 1, 2,
 * 3", SC'11): 10 rounds of two 32x32->64 multiplies with a Weyl key
 * schedule. */
void orc_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                uint32_t k0, uint32_t k1, uint32_t out[4]) {
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

static inline float u01(uint32_t b) { return (float)(b >> 8) * 5.9604644775390625e-8f; }
static inline float u01_open(uint32_t b) {
  return (float)((b >> 8) + 1u) * 5.9604644775390625e-8f;
}
/* Box-Muller in double, rounded once to fp32 (same form in the product). */
static inline float box_muller(uint32_t a, uint32_t b) {
  double u1 = (double)u01_open(a), u2 = (double)u01(b);
  return (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2));
}

void orc_brownian(float *pos, int n, float strength, float L, uint64_t seed,
                  int frame, int mode) {
  const float half_l = L * 0.5f;
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  for (int i = 0; i < n; ++i) {
    uint32_t r0[4], r1[4];
    orc_philox((uint32_t)i, (uint32_t)frame, 0u, 0u, k0, k1, r0);
    float d[3];
    if (mode == 0) {
      for (int a = 0; a < 3; ++a) d[a] = ((u01(r0[a]) - 0.5f) * 2.0f) * strength;
    } else {
      orc_philox((uint32_t)i, (uint32_t)frame, 1u, 0u, k0, k1, r1);
      d[0] = box_muller(r0[0], r0[1]) * strength;
      d[1] = box_muller(r0[2], r0[3]) * strength;
      d[2] = box_muller(r1[0], r1[1]) * strength;
    }
    for (int a = 0; a < 3; ++a) {
      float p = pos[3 * i + a] + d[a];
      if (p < -half_l)
        p = p + L;
      else if (p >= half_l)
        p = p - L;
      pos[3 * i + a] = p;
    }
  }
}

void orc_ou_jitter(float *pos, float *v_jit, const float *rad, int n,
                   float mean_r, float dt, float jitter_rms, float jitter_tau,
                   float drift_cap_frac, float L,
                   uint64_t seed, int frame) {
  const float tau = jitter_tau > 1e-6f ? jitter_tau : 1e-6f;
  const float sigma = jitter_rms * mean_r * sqrtf(2.0f / tau);
  const float sdt = sigma * sqrtf(dt);
  const float capf = drift_cap_frac;
  const float inv_l = (float)(1.0 / (double)L);
  const uint32_t k0 = (uint32_t)seed ^ 0x5bd1e995u, k1 = (uint32_t)(seed >> 32);
  for (int i = 0; i < n; ++i) {
    uint32_t r[4];
    orc_philox((uint32_t)i, (uint32_t)frame, 2u, 0u, k0, k1, r);
    float v[3];
    for (int a = 0; a < 3; ++a) {
      float x = v_jit[3 * i + a];
      float rr = u01(r[a]) - 0.5f;
      v[a] = x + ((-x / tau) * dt + sdt * rr);
    }
    float cap = capf * (rad[i] + rad[i]);
    float vv = v[0] * v[0] + v[1] * v[1];
    vv = vv + v[2] * v[2];
    float len = sqrtf(vv) * dt;
    if (len > cap) {
      float s = (cap / fmaxf(1e-6f, len)) / dt;
      v[0] *= s;
      v[1] *= s;
      v[2] *= s;
    }
    for (int a = 0; a < 3; ++a) {
      v_jit[3 * i + a] = v[a];
      pos[3 * i + a] = gwrap(pos[3 * i + a] + v[a] * dt, L, inv_l);
    }
  }
}

void orc_xsph(const float *pos, float *vel, float *vel_tmp, int n, int G,
              float L, float lim_sq, float eps, const int32_t *cs,
              const int32_t *cc, const int32_t *ci) {
  const float half_l = (float)(0.5 * (double)L), inv_l = (float)(1.0 / (double)L);
  const float lim = lim_sq;
  for (int i = 0; i < n; ++i) {
    const float *pi_p = pos + 3 * i;
    int mc[3];
    centre_cell(pi_p, G, L, half_l, inv_l, mc);
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    int cnt = 0;
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = -1; dz <= 1; ++dz) {
          int id = wrapi(mc[0] + dx, G) * G * G + wrapi(mc[1] + dy, G) * G + wrapi(mc[2] + dz, G);
          for (int k = cs[id]; k < cs[id] + cc[id]; ++k) {
            int j = ci[k];
            if (j == i) continue;
            const float *pj = pos + 3 * j;
            float ddx = gwrap(pj[0] - pi_p[0], L, inv_l);
            float ddy = gwrap(pj[1] - pi_p[1], L, inv_l);
            float ddz = gwrap(pj[2] - pi_p[2], L, inv_l);
            float d2 = ddx * ddx + ddy * ddy;
            d2 = d2 + ddz * ddz;
            if (d2 < lim) {
              sx += vel[3 * j];
              sy += vel[3 * j + 1];
              sz += vel[3 * j + 2];
              cnt += 1;
            }
          }
        }
    float ax = vel[3 * i], ay = vel[3 * i + 1], az = vel[3 * i + 2];
    if (cnt > 0) {
      float fc = (float)cnt;
      ax = sx / fc;
      ay = sy / fc;
      az = sz / fc;
    }
    vel_tmp[3 * i] = (1.0f - eps) * vel[3 * i] + eps * ax;
    vel_tmp[3 * i + 1] = (1.0f - eps) * vel[3 * i + 1] + eps * ay;
    vel_tmp[3 * i + 2] = (1.0f - eps) * vel[3 * i + 2] + eps * az;
  }
  memcpy(vel, vel_tmp, (size_t)n * 3 * sizeof(float));
}

void orc_topo_candidates(const float *pos, const float *rad, int n, int G,
                         float L, float mult, float r_max, int cap,
                         const int32_t *cs, const int32_t *cc,
                         const int32_t *ci, int32_t *pairs, int32_t *count,
                         int32_t *truncated) {
  const float half_l = (float)(0.5 * (double)L), inv_l = (float)(1.0 / (double)L);
  (void)r_max; /* topology.py:98-99 computes an unused bound */
  for (int i = 0; i < n; ++i) {
    count[i] = 0;
    truncated[i] = 0;
    int mc[3];
    centre_cell(pos + 3 * i, G, L, half_l, inv_l, mc);
    for (int d = 0; d < 3; ++d) mc[d] = wrapi(mc[d], G);
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = -1; dz <= 1; ++dz) {
          int id = wrapi(mc[0] + dx, G) * G * G + wrapi(mc[1] + dy, G) * G + wrapi(mc[2] + dz, G);
          for (int k = cs[id]; k < cs[id] + cc[id]; ++k) {
            int j = ci[k];
            if (j > i) {
              const float *pi_p = pos + 3 * i, *pj = pos + 3 * j;
              float ddx = gwrap(pj[0] - pi_p[0], L, inv_l);
              float ddy = gwrap(pj[1] - pi_p[1], L, inv_l);
              float ddz = gwrap(pj[2] - pi_p[2], L, inv_l);
              float d2 = ddx * ddx + ddy * ddy;
              d2 = d2 + ddz * ddz;
              float th = mult * (rad[i] + rad[j]);
              if (d2 <= th * th) {
                if (count[i] < cap) {
                  pairs[(int64_t)i * cap + count[i]] = j;
                  count[i] += 1;
                } else {
                  truncated[i] = 1;
                  break; /* topology.py:131: leaves this cell only */
                }
              }
            }
          }
        }
  }
}

/* ---- topology.py:140-504: Gabriel / kNN degrees, EMA, radius rule ---- */

/* topology.py:188 and :326-333 pair hash (int32 arithmetic wraps). */
static inline int32_t topo_hash(int i, int j, int frame) {
  uint32_t h = ((uint32_t)i * 73856093u) ^ ((uint32_t)j * 19349663u) ^
               ((uint32_t)frame * 83492791u);
  return (int32_t)h;
}

/* grid.py:46 wrapP applied to one component. */
static inline void wrap3(float v[3], float L, float inv_l) {
  for (int d = 0; d < 3; ++d) v[d] = gwrap(v[d], L, inv_l);
}

/* topology.py:198-252 witness test of pair (i, j): 1 when no other particle
 * lies strictly inside the diameter sphere (scaled epsilon). */
static int gabriel_pair(const float *pos, int i, int j, int G, float L, float cell_size,
                        const int32_t *cs, const int32_t *cc, const int32_t *ci) {
  const float half_l = (float)(0.5 * (double)L), inv_l = (float)(1.0 / (double)L);
  const float *pi_p = pos + 3 * i, *pj = pos + 3 * j;
  float d[3] = {pj[0] - pi_p[0], pj[1] - pi_p[1], pj[2] - pi_p[2]};
  wrap3(d, L, inv_l);
  float d2 = d[0] * d[0] + d[1] * d[1];
  d2 = d2 + d[2] * d[2];
  if (d2 < 1e-16f) return -1; /* skipped pair */
  float mid[3];
  for (int a = 0; a < 3; ++a) mid[a] = pi_p[a] + 0.5f * d[a];
  wrap3(mid, L, inv_l);
  const float r_sq = 0.25f * d2;
  int mc[3];
  for (int a = 0; a < 3; ++a) {
    float q = (mid[a] + half_l) * inv_l;
    mc[a] = wrapi((int)(q * (float)G), G);
  }
  const float r_d = sqrtf(d2) * 0.5f;
  int sc = (int)ceilf(r_d / cell_size);
  sc = sc < 1 ? 1 : (sc > 3 ? 3 : sc);
  if (sc > G / 2) sc = G / 2;
  const float eps = fmaxf(1e-12f, 1e-6f * r_sq);
  for (int dx = -sc; dx <= sc; ++dx)
    for (int dy = -sc; dy <= sc; ++dy)
      for (int dz = -sc; dz <= sc; ++dz) {
        int id = wrapi(mc[0] + dx, G) * G * G + wrapi(mc[1] + dy, G) * G + wrapi(mc[2] + dz, G);
        for (int l = cs[id]; l < cs[id] + cc[id]; ++l) {
          int k = ci[l];
          if (k == i || k == j) continue;
          float e[3] = {mid[0] - pos[3 * k], mid[1] - pos[3 * k + 1], mid[2] - pos[3 * k + 2]};
          wrap3(e, L, inv_l);
          float dk = e[0] * e[0] + e[1] * e[1];
          dk = dk + e[2] * e[2];
          if (dk < r_sq - eps) return 0;
        }
      }
  return 1;
}

void orc_topo_gabriel(const float *pos, int n, int G, float L, float cell_size,
                      const int32_t *cs, const int32_t *cc, const int32_t *ci,
                      const int32_t *pairs, const int32_t *pair_count, int cap, int frame,
                      int stride, int early_exit, float deg_high, const float *deg_ema,
                      int32_t *deg) {
  for (int i = 0; i < n; ++i) deg[i] = 0;
  for (int i = 0; i < n; ++i) {
    int pc = pair_count[i] < cap ? pair_count[i] : cap;
    for (int p = 0; p < pc; ++p) {
      int j = pairs[(int64_t)i * cap + p];
      if (stride > 1 && topo_hash(i, j, frame) % stride != 0) continue;
      if (early_exit && deg_ema[i] >= deg_high && deg_ema[j] >= deg_high) continue;
      if (gabriel_pair(pos, i, j, G, L, cell_size, cs, cc, ci) == 1 && j > i) {
        deg[i] += 1;
        deg[j] += 1;
      }
    }
  }
}

void orc_topo_gabriel_batched(const float *pos, const float *rad, int n, int G, float L,
                              float cell_size, float mult, int cap, const int32_t *cs,
                              const int32_t *cc, const int32_t *ci, int batch_idx,
                              int num_batches, int q, int frame, int32_t *pairs,
                              int32_t *pair_count, int32_t *truncated, int32_t *deg) {
  const float half_l = (float)(0.5 * (double)L), inv_l = (float)(1.0 / (double)L);
  for (int i = 0; i < n; ++i) {
    if (i % num_batches != batch_idx) continue;
    int count = 0;
    truncated[i] = 0;
    int mc[3];
    centre_cell(pos + 3 * i, G, L, half_l, inv_l, mc);
    for (int d = 0; d < 3; ++d) mc[d] = wrapi(mc[d], G);
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = -1; dz <= 1; ++dz) {
          int id = wrapi(mc[0] + dx, G) * G * G + wrapi(mc[1] + dy, G) * G + wrapi(mc[2] + dz, G);
          for (int k = cs[id]; k < cs[id] + cc[id]; ++k) {
            int j = ci[k];
            if (j <= i) continue;
            if (topo_hash(i, j, frame) % q != 0) continue;
            float e[3] = {pos[3 * j] - pos[3 * i], pos[3 * j + 1] - pos[3 * i + 1],
                          pos[3 * j + 2] - pos[3 * i + 2]};
            wrap3(e, L, inv_l);
            float d2 = e[0] * e[0] + e[1] * e[1];
            d2 = d2 + e[2] * e[2];
            float th = mult * (rad[i] + rad[j]);
            if (d2 <= th * th) {
              int cur = count++;
              if (cur < cap)
                pairs[(int64_t)i * cap + cur] = j;
              else
                truncated[i] = 1;
            }
          }
        }
    /* topology.py:385 counts past the list; only stored pairs are tested */
    pair_count[i] = count;
    int pc = count < cap ? count : cap;
    for (int p = 0; p < pc; ++p) {
      int j = pairs[(int64_t)i * cap + p];
      if (gabriel_pair(pos, i, j, G, L, cell_size, cs, cc, ci) == 1) {
        deg[i] += 1;
        deg[j] += 1;
      }
    }
  }
}

void orc_topo_knn(const float *pos, const float *rad, int n, int G, float L, float mult, int k,
                  const int32_t *cs, const int32_t *cc, const int32_t *ci, int32_t *deg) {
  const float half_l = (float)(0.5 * (double)L), inv_l = (float)(1.0 / (double)L);
  for (int i = 0; i < n; ++i) {
    int cnt = 0;
    float r = mult * rad[i];
    float th = r * r;
    int mc[3];
    centre_cell(pos + 3 * i, G, L, half_l, inv_l, mc);
    for (int d = 0; d < 3; ++d) mc[d] = wrapi(mc[d], G);
    for (int dx = -1; dx <= 1 && cnt < k; ++dx)
      for (int dy = -1; dy <= 1 && cnt < k; ++dy)
        for (int dz = -1; dz <= 1 && cnt < k; ++dz) {
          int id = wrapi(mc[0] + dx, G) * G * G + wrapi(mc[1] + dy, G) * G + wrapi(mc[2] + dz, G);
          for (int l = cs[id]; l < cs[id] + cc[id] && cnt < k; ++l) {
            int j = ci[l];
            if (j == i) continue;
            float e[3] = {pos[3 * j] - pos[3 * i], pos[3 * j + 1] - pos[3 * i + 1],
                          pos[3 * j + 2] - pos[3 * i + 2]};
            wrap3(e, L, inv_l);
            float d2 = e[0] * e[0] + e[1] * e[1];
            d2 = d2 + e[2] * e[2];
            if (d2 <= th) cnt += 1;
          }
        }
    deg[i] = cnt < k ? cnt : k;
  }
}

/* topology.py:266-283 update_topo_ema and :449-455 commit_topo_to_ema. */
void orc_topo_ema(const int32_t *deg, float *ema, int n, float alpha) {
  for (int i = 0; i < n; ++i) ema[i] = alpha * (float)deg[i] + (1.0f - alpha) * ema[i];
}

/* topology.py:287-322 update_radii_topological. */
void orc_topo_radii(float *rad, const float *deg_eff, int n, float deg_low, float deg_high,
                    float gain_grow, float gain_shrink, float r_min, float r_max) {
  for (int i = 0; i < n; ++i) {
    float d = deg_eff[i];
    if (d < deg_low)
      rad[i] *= (1.0f + gain_grow);
    else if (d > deg_high)
      rad[i] *= (1.0f - gain_shrink);
    rad[i] = fmaxf(r_min, fminf(r_max, rad[i]));
  }
}
