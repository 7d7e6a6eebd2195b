"""CPU restatement of one foam frame (run.py:700-1261 order) over liboracle.

TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline).  The state and
the configuration fields match the product driver so a single frame can be
compared from identical state.  PBD uses the Jacobi form (every particle
reads the pass-start positions); ``gauss_seidel=True`` selects the literal
index-order update of the reference's serial semantics.
"""
import math

import numpy as np

from . import oracle as O

MAXN = 32


def f32(x):
    return float(np.float32(x))


class OracleFoam:
    def __init__(self, cfg, pos, rad, gauss_seidel=False):
        self.cfg = cfg
        n = self.n = int(cfg.n)
        self.pos = np.array(pos, np.float32, copy=True).reshape(n, 3)
        self.rad = np.array(rad, np.float32, copy=True).reshape(n)
        self.rad_target = self.rad.copy()
        self.lmd = np.zeros(n, np.float32)
        self.v_jit = np.zeros((n, 3), np.float32)
        self.G = cfg.hash_res()
        self.face_ids = np.full((n, MAXN), -1, np.int32)
        self.face_counts = np.zeros((n, MAXN), np.int32)
        self.fsc = np.zeros(n, np.int32)
        self.fsc_ema = np.zeros(n, np.float32)
        self.neighbor_list = np.full((n, MAXN), -1, np.int32)
        self.neighbor_count = np.zeros(n, np.int32)
        self.used_count = np.zeros(n, np.int32)
        self.prev_scores = np.zeros((n, MAXN), np.float32)
        self.frame = 0
        self.settle = 0
        self.gauss_seidel = gauss_seidel
        self.grid = None
        self.labels = None
        self.filtered = None

    def rebuild_grid(self):
        c = self.cfg
        self.pos, self.cc, self.cs, self.ci = O.grid_build(self.pos, self.n, self.G, c.L)

    def run_jfa(self):
        c = self.cfg
        voxel = c.voxel()
        self.labels, passes, early, hist = O.jfa_run(
            self.pos, self.rad, self.n, c.res, c.L, voxel, c.passes(), c.early_exit,
            c.th_pass_delta, c.min_passes, c.beta)
        self.jfa_info = (passes, early, hist)
        self.filtered = O.majority(self.labels, c.res)
        self.face_ids, self.face_counts = O.accumulate_faces(self.filtered, c.res, self.n)
        (self.prev_scores, self.neighbor_list, self.neighbor_count, self.used_count,
         self.fsc) = O.finalize_faces(self.n, self.rad, voxel, self.face_ids, self.face_counts,
                                      self.prev_scores)
        self.fsc_ema = O.fsc_ema(self.n, c.ema_alpha, self.fsc, self.fsc_ema)

    def step(self):
        c = self.cfg
        f = self.frame
        n = self.n
        self.rebuild_grid()
        max_depth, self.lmd = O.max_overlap(self.pos, self.rad, n, self.G, c.L, self.cs,
                                            self.cc, self.ci)
        passes = max(c.pbd_base, min(c.pbd_max,
                                     int(c.pbd_base + c.pbd_scale * (max_depth / (0.2 * c.r_max)))))
        if self.settle > 0:
            back = max(0.25, min(1.0, (c.contact_tol - max_depth) / c.contact_tol))
            self.rad = O.nudge_radii(self.rad, self.rad_target, self.lmd, n, c.alpha(),
                                     c.max_step_pct, c.r_min, c.r_max, back,
                                     c.backpressure_mode, c.contact_tol)
            self.settle -= 1
        if c.pressure:
            for _ in range(c.pressure_iters):
                self.rad, _, _, _ = O.pressure(n, f, self.rad, self.face_ids, self.fsc,
                                               c.pressure_k, c.pressure_exp, c.pressure_cap,
                                               c.pressure_max_nei, c.r_min, c.r_max)
        if c.brownian:
            self.pos = O.brownian(self.pos, n, c.brownian_strength, c.L, c.seed, f,
                                  c.brownian_mode)
        for _ in range(passes):
            self.rebuild_grid()
            args = (self.rad, n, self.G, c.L, c.gap_factor(), c.max_disp_frac, c.disp_mult,
                    self.cs, self.cc, self.ci)
            if self.gauss_seidel:
                self.pos = O.pbd_gauss_seidel(self.pos, *args)
            else:
                self.pos = O.pbd_jacobi(self.pos, *args)
        if c.ou_jitter:
            mean_r = f32(np.mean(self.rad, dtype=np.float64))
            self.pos, self.v_jit = O.ou_jitter(self.pos, self.v_jit, self.rad, n, mean_r,
                                               f32(0.016), f32(0.10), f32(1.0),
                                               f32(0.20 * c.gap_fraction), c.L, c.seed, f)
        self.rebuild_grid()
        ran = c.jfa_runs(f)
        if ran:
            self.run_jfa()
            f0 = float(np.sum(self.fsc == 0)) / float(max(1, n))
            growth = not ((f < c.warmstart_frames) or (f0 > c.fsc_zero_rate))
        else:
            growth = True
        if growth:
            self.rad_target = O.fsc_targets(self.rad, self.fsc, self.fsc_ema, n, c.f_low,
                                            c.f_high, c.growth_pct, c.r_min, c.r_max,
                                            c.deadband)
            self.settle = c.adjustment_frames
        self.frame += 1
        return {"frame": f, "pbd_passes": passes, "max_depth": max_depth, "jfa": ran,
                "growth": growth}

    STATE = ("pos", "rad", "rad_target", "lmd", "v_jit", "face_ids", "face_counts", "fsc",
             "fsc_ema", "neighbor_list", "neighbor_count", "used_count", "prev_scores")

    def snapshot(self):
        out = {k: np.array(getattr(self, k), copy=True) for k in self.STATE}
        out["frame"] = np.int64(self.frame)
        out["settle"] = np.int64(self.settle)
        return out

    def restore(self, snap):
        for k in self.STATE:
            setattr(self, k, np.array(snap[k], copy=True))
        self.frame = int(snap["frame"])
        self.settle = int(snap["settle"])
