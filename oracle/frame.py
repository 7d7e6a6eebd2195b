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
        # temporal-skip scheduler state (run.py:165-172)
        self.pos_prev = np.zeros((n, 3), np.float32)
        self.rad_prev = np.zeros(n, np.float32)
        self.sched_frames = 0
        self.since = 0
        self.have_prev = False
        self.last_fsc_mu = None
        self.fsc_mu_pending = False
        self.lockout = 0
        self.reason = None
        self.jfa_res = cfg.res

    def rebuild_grid(self):
        c = self.cfg
        self.pos, self.cc, self.cs, self.ci = O.grid_build(self.pos, self.n, self.G, c.L)

    def sched_stats(self, have_prev):
        """run.py:913-948 signals, with the sums in fp64 (out[0..5] as the
        library's fos_sched_stats documents them); moves the current radii
        and positions into the previous-value arrays."""
        c, n = self.cfg, self.n
        out = np.zeros(6)
        if have_prev:
            dr = np.abs(self.rad - self.rad_prev).astype(np.float64)
            out[0] = dr.sum() / n
            out[1] = math.sqrt(((dr - out[0]) ** 2).sum() / n)
            d = self.pos - self.pos_prev
            half = np.float32(c.L) * np.float32(0.5)
            d = np.where(np.abs(d) > half, d - np.copysign(np.float32(c.L), d), d)
            d = d.astype(np.float32)
            s2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
            out[2] = float(np.sqrt(s2.astype(np.float32)).max())
        out[3] = self.rad.astype(np.float64).sum() / n
        out[4] = self.fsc.astype(np.float64).sum() / n
        out[5] = float(np.sum((self.fsc < c.f_low) | (self.fsc > c.f_high)))
        self.rad_prev = self.rad.copy()
        self.pos_prev = self.pos.copy()
        return out

    def schedule(self, f):
        """The JFA decision (cadence, or run.py:902-1018) and the resolution
        (run.py:1049-1068 when adaptive)."""
        c = self.cfg
        temporal = c.jfa_schedule == "temporal"
        res = c.res
        o = None
        if temporal:
            o = self.sched_stats(self.have_prev)
            r_mean = o[3]
            self.sched_frames += 1
            self.since += 1
            sig = o[1] / max(r_mean, 1e-9) if self.have_prev else 999.0
            vmax = o[2] if self.have_prev else 999.0
            self.have_prev = True
            if self.fsc_mu_pending:
                self.last_fsc_mu, self.fsc_mu_pending = o[4], False
            dmu = abs(o[4] - self.last_fsc_mu) if self.last_fsc_mu is not None else 999.0
            if self.lockout > 0:
                self.lockout -= 1
            warm = self.sched_frames <= c.skip_warmstart
            watchdog = self.since >= c.max_skip
            band = self.lockout > 0
            activity = sig > c.th_rsigma or vmax > c.th_margin * r_mean or dmu > c.th_fsc_drift
            oob = self.since > 1 and o[5] / max(1, self.n) > c.oob_frac
            run = warm or watchdog or band or activity or oob
            self.reason = ("warm-start" if warm else "band-change" if band else
                           "watchdog" if watchdog else "out-of-band-guard" if oob else
                           "activity" if activity else "skip")
            if run:
                self.since = 0
                self.fsc_mu_pending = True
        else:
            run = c.jfa_runs(f)
            self.reason = "cadence" if run else "skip"
        if run and c.adaptive_res:
            if o is None:
                o = self.sched_stats(False)
            r = int(round(float(c.L) * float(c.voxels_per_diameter) / (2.0 * o[3])))
            res = max(c.res_min, min(r, c.res_max))
        return run, res

    def set_band(self, f_low, f_high):
        c = self.cfg
        if (f_low, f_high) != (c.f_low, c.f_high):
            c.f_low, c.f_high = f_low, f_high
            if c.jfa_schedule == "temporal":
                self.lockout = c.band_lockout

    def run_jfa(self, res=None):
        c = self.cfg
        res = c.res if res is None else res
        voxel = c.voxel() if res == c.res else f32(c.L / res)
        npass = c.passes() if res == c.res else int(math.ceil(math.log2(res))) + 1
        self.jfa_res = res
        self.labels, passes, early, hist = O.jfa_run(
            self.pos, self.rad, self.n, res, c.L, voxel, npass, c.early_exit,
            c.th_pass_delta, c.min_passes, c.beta)
        self.jfa_info = (passes, early, hist)
        self.filtered = O.majority(self.labels, res)
        self.face_ids, self.face_counts = O.accumulate_faces(self.filtered, res, self.n)
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
        ran, res = self.schedule(f)
        if ran:
            self.run_jfa(res)
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
                "growth": growth, "jfa_reason": self.reason,
                "jfa_res": self.jfa_res if ran else None}

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
