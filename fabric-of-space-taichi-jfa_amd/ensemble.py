"""Batched ensembles: B independent foams of uneven size advanced together.

No reference counterpart (the reference runs one foam per process).  The
foams' particles, JFA labels and hash cells are concatenated (foam b owns
particles [p0_b, p0_b + n_b), labels [v0_b, v0_b + res_b^3) and cells
[c0_b, c0_b + G_b^3)); particle ids are batch-global.  Every phase of the
frame is one launch over all foams (``fos_batch_*``), so a batch of small
foams fills the GPU where one small foam cannot.

Each foam follows the semantics of ``foam.FoamSim`` on its own inputs: the
per-foam results equal B separate FoamSim runs (tests/test_gpu_batch.py).
"""
import ctypes
import math

import numpy as np
import torch

import fos_lib as fl
import foam_config as fc

MAXN = 32
SLOTS = 16          # pass counter slots per foam (FOS_BATCH_PASSES)
FIXED_PAIRS = 16    # finalize round pairs in a captured frame

# settings every foam of a batch shares (the per-foam ones are n, res, the
# radius bounds, the hash grid, the FSC band and the seed)
SHARED = ("L", "num_passes", "early_exit", "th_pass_delta", "min_passes", "beta", "ema_alpha",
          "pbd_base", "pbd_max", "pbd_scale", "gap_fraction", "max_disp_frac", "disp_mult",
          "growth_pct", "adjustment_frames", "max_step_pct", "deadband", "backpressure_mode",
          "contact_tol", "warmstart_frames", "fsc_zero_rate", "pressure", "pressure_k",
          "pressure_exp", "pressure_cap", "pressure_max_nei", "pressure_iters", "brownian",
          "brownian_mode", "brownian_strength", "ou_jitter", "jfa_every", "jfa_warm")


def c4_configs(B=256, seed=0, n_min=500, n_max=20000):
    """The batched-ensemble workload: N log-uniform in [n_min, n_max], radii
    U[0.6, 1.2] * (0.5 / N)^(1/3), JFA edge round(5 N^(1/3)) clamped to
    [48, 160], FSC band from {10-14, 12-16, 14-18}, independent seeds; the
    dynamics of the C2 workload otherwise."""
    rng = np.random.default_rng(seed)
    bands = ((10, 14), (12, 16), (14, 18))
    out = []
    for b in range(B):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        s = (0.5 / n) ** (1.0 / 3.0)
        res = int(min(160, max(48, round(5.0 * n ** (1.0 / 3.0)))))
        lo, hi = bands[int(rng.integers(0, 3))]
        cfg = fc.preset("c2")
        cfg.ou_jitter = False          # the batch has no OU jitter kernel
        cfg.name = f"c4[{b}]"
        cfg.n, cfg.res = n, res
        cfg.r_min, cfg.r_max = fc.f32(0.6 * s * 0.5), fc.f32(1.2 * s * 1.0)
        cfg.f_low, cfg.f_high = lo, hi
        cfg.seed = int(rng.integers(1, 2 ** 31 - 1))
        cfg.frames = 100
        out.append(cfg)
    return out


def c4_inputs(cfg, b):
    """Seeded uniform inputs of foam b: radii U[0.6, 1.2] * (0.5 / N)^(1/3)."""
    s = (0.5 / cfg.n) ** (1.0 / 3.0)
    rng = np.random.default_rng(1000003 * (b + 1) + cfg.n)
    pos = (rng.random((cfg.n, 3), dtype=np.float64) - 0.5).astype(np.float32) * np.float32(cfg.L)
    rad = (rng.uniform(0.6, 1.2, cfg.n) * s).astype(np.float32)
    return pos, rad


class Ensemble:
    """B foams with per-foam configs (only n, res, r_min/r_max, grid_res,
    f_low/f_high and seed may differ) and per-foam inputs."""

    def __init__(self, cfgs, inputs, device="cuda"):
        fl.require_gpu()
        self.cfgs = list(cfgs)
        B = len(self.cfgs)
        c0 = self.cfgs[0]
        for c in self.cfgs[1:]:
            for k in SHARED:
                if getattr(c, k) != getattr(c0, k):
                    raise ValueError(f"foams of a batch must share {k}")
        self.cfg = c0
        self.B = B
        dev = torch.device(device)
        ns = [int(c.n) for c in self.cfgs]
        self.n_of = np.array(ns, np.int64)
        self.p0 = np.concatenate([[0], np.cumsum(self.n_of)]).astype(np.int64)
        n = int(self.p0[-1])
        self.n = n
        self.res_of = np.array([int(c.res) for c in self.cfgs], np.int64)
        self.G_of = np.array([c.hash_res() for c in self.cfgs], np.int64)
        self.v0 = np.concatenate([[0], np.cumsum(self.res_of ** 3)]).astype(np.int64)
        self.c0 = np.concatenate([[0], np.cumsum(self.G_of ** 3)]).astype(np.int64)
        I = ctypes.c_int
        arr = lambda v, t: (t * B)(*v)
        h = ctypes.c_void_p()
        fl.call("fos_batch_create", ctypes.byref(h), B, arr(ns, I),
                arr([int(c.res) for c in self.cfgs], I),
                arr([int(c.passes()) for c in self.cfgs], I), arr([int(g) for g in self.G_of], I),
                arr([c.r_min for c in self.cfgs], ctypes.c_float),
                arr([c.r_max for c in self.cfgs], ctypes.c_float),
                arr([float(c.f_low) for c in self.cfgs], ctypes.c_float),
                arr([float(c.f_high) for c in self.cfgs], ctypes.c_float),
                arr([int(c.seed) & 0xFFFFFFFF for c in self.cfgs], ctypes.c_uint), float(c0.L))
        self.handle = h
        pos = np.concatenate([np.asarray(p, np.float32).reshape(-1, 3) for p, _ in inputs])
        rad = np.concatenate([np.asarray(r, np.float32).reshape(-1) for _, r in inputs])
        assert pos.shape[0] == n and rad.shape[0] == n
        f32, i32 = torch.float32, torch.int32
        self.pos = torch.as_tensor(pos, device=dev).contiguous()
        self.rad = torch.as_tensor(rad, device=dev).contiguous()
        V = int(self.v0[-1])
        self.grid = torch.full((V,), -1, dtype=i32, device=dev)
        self.temp = torch.full((V,), -1, dtype=i32, device=dev)
        self.filtered = torch.full((V,), -1, dtype=i32, device=dev)
        self.face_ids = torch.full((n, MAXN), -1, dtype=i32, device=dev)
        self.face_counts = torch.zeros(n, MAXN, dtype=i32, device=dev)
        self.fsc = torch.zeros(n, dtype=i32, device=dev)
        self.fsc_ema = torch.zeros(n, dtype=f32, device=dev)
        self.neighbor_list = torch.full((n, MAXN), -1, dtype=i32, device=dev)
        self.neighbor_count = torch.zeros(n, dtype=i32, device=dev)
        self.used_count = torch.zeros(n, dtype=i32, device=dev)
        self.prev_scores = torch.zeros(n, MAXN, dtype=f32, device=dev)
        self.info = torch.zeros(B, 2 + SLOTS, dtype=i32, device=dev)
        C = int(self.c0[-1])
        self.cell_count = torch.zeros(C, dtype=i32, device=dev)
        self.cell_start = torch.zeros(C, dtype=i32, device=dev)
        self.cell_indices = torch.zeros(n, dtype=i32, device=dev)
        self.cell_pr = torch.zeros(n, 4, dtype=f32, device=dev)   # slot (position, radius)
        self.pos_tmp = torch.empty_like(self.pos)
        self.rad_target = self.rad.clone()
        self.lmd = torch.zeros(n, dtype=f32, device=dev)
        self.delta_r = torch.zeros(n, dtype=f32, device=dev)
        self.maxd = torch.zeros(B, dtype=f32, device=dev)
        self.zeros = torch.zeros(B, dtype=i32, device=dev)
        # per-foam controls of a frame, uploaded by the host: [0] PBD passes,
        # [1] settle active, [2] growth active (int); back factor (float)
        self.ctl = torch.zeros(3, B, dtype=i32, device=dev)
        self.back = torch.zeros(B, dtype=f32, device=dev)
        # its own context: a captured frame holds pointers into its scratch
        self.ctx = fl.Context(n, int(self.res_of.max()), C)
        self.frame = 0
        self.tiles = [0] * B
        self.tile_choice = "rows"
        self.settle = np.zeros(B, np.int64)
        self.growth = np.ones(B, bool)
        self.last = {}

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                fl.call("fos_batch_destroy", h)
            except Exception:
                pass
            self.handle = None

    # -- JFA tile shapes ------------------------------------------------------------
    def set_tiles(self, kinds):
        """JFA pass tile per foam: 0 rows (4 z-rows of 64 voxels per
        workgroup), 1 flat (256 consecutive voxels of a y-z plane)."""
        self.tiles = [int(k) for k in kinds]
        fl.call("fos_batch_set_tiles", self.handle, (ctypes.c_int * self.B)(*self.tiles))

    def autotune_tiles(self, reps=3):
        """Times this batch's JFA passes under candidate tile assignments --
        all row tiles, all flat tiles, and per foam the kind with fewer idle
        lanes at its grid edge -- keeps the fastest and returns the times
        (ms per pass sequence, best of reps).  Labels are re-seeded before
        each timed run; results never depend on the tiles."""
        def idle_free(r):
            row = r / (64.0 * math.ceil(r / 64.0))
            flat = r * r / (256.0 * math.ceil(r * r / 256.0))
            return 0 if row >= flat else 1
        cands = {"rows": [0] * self.B, "flat": [1] * self.B,
                 "per_edge": [idle_free(int(r)) for r in self.res_of]}
        c, st = self.cfg, fl.stream()
        times = {}
        for name, kinds in cands.items():
            self.set_tiles(kinds)
            best = None
            for _ in range(reps):
                fl.call("fos_batch_jfa_init", self.handle, self.ctx.handle, fl.ptr(self.pos),
                        fl.ptr(self.rad), float(c.L), float(c.beta), fl.ptr(self.grid), None, st)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                fl.call("fos_batch_jfa_passes", self.handle, self.ctx.handle, float(c.L),
                        int(c.early_exit), c.th_pass_delta, c.min_passes, fl.ptr(self.grid),
                        fl.ptr(self.temp), fl.ptr(self.info), st)
                e1.record()
                e1.synchronize()
                ms = e0.elapsed_time(e1)
                best = ms if best is None else min(best, ms)
            times[name] = round(best, 3)
        pick = min(times, key=times.get)
        self.set_tiles(cands[pick])
        self.tile_choice = pick
        return times

    def foam(self, b, t):
        """Foam b's rows of a batch tensor indexed by particle."""
        return t[int(self.p0[b]):int(self.p0[b + 1])]

    def labels(self, b, t=None):
        """Foam b's label block (res^3, batch-global ids) of grid or t."""
        t = self.grid if t is None else t
        return t[int(self.v0[b]):int(self.v0[b + 1])]

    def run_jfa(self, graph=False):
        """jfa.run_jfa for every foam (jfa.py:1078-1132), batched.  graph:
        no host round trips (the flags are read after the frame)."""
        c, ctx, st, h = self.cfg, self.ctx.handle, fl.stream(), self.handle
        fl.call("fos_batch_jfa_init", h, ctx, fl.ptr(self.pos), fl.ptr(self.rad), float(c.L),
                float(c.beta), fl.ptr(self.grid), None, st)
        fl.call("fos_batch_jfa_passes", h, ctx, float(c.L), int(c.early_exit), c.th_pass_delta,
                c.min_passes, fl.ptr(self.grid), fl.ptr(self.temp), fl.ptr(self.info), st)
        fl.call("fos_batch_majority", h, fl.ptr(self.grid), fl.ptr(self.filtered), st)
        fl.call("fos_batch_face_accumulate", h, ctx, fl.ptr(self.filtered),
                fl.ptr(self.face_ids), fl.ptr(self.face_counts), int(not graph), st)
        fl.call("fos_batch_finalize_faces", h, ctx, fl.ptr(self.rad), fl.ptr(self.face_ids),
                fl.ptr(self.face_counts), fl.ptr(self.prev_scores), fl.ptr(self.neighbor_list),
                fl.ptr(self.neighbor_count), fl.ptr(self.used_count), fl.ptr(self.fsc),
                FIXED_PAIRS if graph else 0, st)
        fl.call("fos_fsc_ema", self.n, c.ema_alpha, fl.ptr(self.fsc), fl.ptr(self.fsc_ema), st)

    # -- the frame (foam.FoamSim.step for every foam) -----------------------------
    def rebuild_grid(self, passes=None, p=0):
        c = self.cfg
        fl.call("fos_batch_grid_build", self.handle, self.ctx.handle, fl.ptr(self.pos), float(c.L),
                fl.ptr(self.cell_count), fl.ptr(self.cell_start), fl.ptr(self.cell_indices),
                None if passes is None else fl.ptr(passes), int(p), fl.ptr(self.rad),
                fl.ptr(self.cell_pr), fl.stream())

    def step(self):
        """One frame of every foam, in the run.py:700-1261 order (see
        foam.FoamSim.step).  The host reads the foams' overlap maxima (their
        PBD pass counts) and, on JFA frames, their zero-FSC counts."""
        c, f, st, h, B = self.cfg, self.frame, fl.stream(), self.handle, self.B
        if getattr(self, "graphs", None):
            # an eager frame may regrow the face-row tables a graph points into
            self.settle = self.settle_dev.cpu().numpy().astype(np.int64)
            self.graphs = None
        if c.ou_jitter:
            raise NotImplementedError("OU jitter is not batched")
        # 0-1
        self.rebuild_grid()
        fl.call("fos_batch_max_overlap", h, fl.ptr(self.pos), fl.ptr(self.rad), float(c.L),
                fl.ptr(self.cell_start), fl.ptr(self.cell_count), fl.ptr(self.cell_indices),
                fl.ptr(self.cell_pr), fl.ptr(self.lmd), fl.ptr(self.maxd), st)
        maxd = self.maxd.cpu().numpy()
        passes = np.array([max(c.pbd_base, min(c.pbd_max, int(c.pbd_base + c.pbd_scale * (
            float(md) / (0.2 * cb.r_max))))) for md, cb in zip(maxd, self.cfgs)], np.int32)
        # 2 settle
        act = self.settle > 0
        if act.any():
            back = np.array([max(0.25, min(1.0, (c.contact_tol - float(md)) / c.contact_tol))
                             for md in maxd], np.float32)
            self.back.copy_(torch.from_numpy(back))
            self.ctl[1].copy_(torch.from_numpy(act.astype(np.int32)))
            fl.call("fos_batch_nudge_radii", h, fl.ptr(self.rad), fl.ptr(self.rad_target),
                    fl.ptr(self.lmd), c.alpha(), c.max_step_pct, c.backpressure_mode,
                    c.contact_tol, fl.ptr(self.back), fl.ptr(self.ctl[1]), st)
            self.settle[act] -= 1
        # 3 pressure
        if c.pressure:
            fl.call("fos_batch_pressure_run", h, self.ctx.handle, f, None, c.pressure_iters,
                    fl.ptr(self.rad), fl.ptr(self.delta_r), fl.ptr(self.face_ids),
                    fl.ptr(self.fsc), c.pressure_k, c.pressure_exp, c.pressure_cap,
                    c.pressure_max_nei, st)
        # 4 Brownian
        if c.brownian:
            fl.call("fos_batch_brownian", h, fl.ptr(self.pos), c.brownian_strength, float(c.L), f,
                    None, c.brownian_mode, st)
        # 5 PBD: a foam with fewer passes sits the later passes out
        self.ctl[0].copy_(torch.from_numpy(passes))
        for p in range(int(passes.max())):
            self.rebuild_grid(self.ctl[0], p)
            fl.call("fos_batch_pbd_pass", h, fl.ptr(self.pos), fl.ptr(self.pos_tmp),
                    fl.ptr(self.rad), float(c.L), c.gap_factor(), c.max_disp_frac, c.disp_mult,
                    fl.ptr(self.cell_start), fl.ptr(self.cell_count), fl.ptr(self.cell_indices),
                    fl.ptr(self.cell_pr), fl.ptr(self.ctl[0]), p, st)
            self.pos, self.pos_tmp = self.pos_tmp, self.pos
        # 7
        self.rebuild_grid()
        # 8 JFA on cadence frames
        ran = c.jfa_runs(f)
        if ran:
            self.run_jfa()
            fl.call("fos_batch_count_zero", h, fl.ptr(self.fsc), fl.ptr(self.zeros), st)
            f0 = self.zeros.cpu().numpy() / np.maximum(1, self.n_of).astype(np.float64)
            self.growth = ~((f < c.warmstart_frames) | (f0 > c.fsc_zero_rate))
        else:
            self.growth = np.ones(B, bool)
        # 9 controller
        if self.growth.any():
            self.ctl[2].copy_(torch.from_numpy(self.growth.astype(np.int32)))
            fl.call("fos_batch_fsc_targets", h, fl.ptr(self.rad), fl.ptr(self.rad_target),
                    fl.ptr(self.fsc), fl.ptr(self.fsc_ema), c.growth_pct, c.deadband,
                    fl.ptr(self.ctl[2]), st)
            self.settle[self.growth] = c.adjustment_frames
        self.last = {"frame": f, "pbd_passes": passes, "max_depth": maxd, "jfa": ran,
                     "growth": self.growth.copy()}
        self.frame += 1
        return self.last

    # -- captured frames (hipGraph) -------------------------------------------
    def _frame_device(self, ran):
        """One frame with every per-foam decision on the device and no host
        round trip: what capture() records.  ran: whether the JFA runs."""
        c, st, h = self.cfg, fl.stream(), self.handle
        self.rebuild_grid()
        fl.call("fos_batch_max_overlap", h, fl.ptr(self.pos), fl.ptr(self.rad), float(c.L),
                fl.ptr(self.cell_start), fl.ptr(self.cell_count), fl.ptr(self.cell_indices),
                fl.ptr(self.cell_pr), fl.ptr(self.lmd), fl.ptr(self.maxd), st)
        fl.call("fos_batch_control_pre", h, fl.ptr(self.maxd), fl.ptr(self.rmax_d), c.pbd_base,
                c.pbd_max, float(c.pbd_scale), float(c.contact_tol), fl.ptr(self.settle_dev),
                fl.ptr(self.ctl[0]), fl.ptr(self.ctl[1]), fl.ptr(self.back), st)
        fl.call("fos_batch_nudge_radii", h, fl.ptr(self.rad), fl.ptr(self.rad_target),
                fl.ptr(self.lmd), c.alpha(), c.max_step_pct, c.backpressure_mode, c.contact_tol,
                fl.ptr(self.back), fl.ptr(self.ctl[1]), st)
        if c.pressure:
            fl.call("fos_batch_pressure_run", h, self.ctx.handle, 0, fl.ptr(self.frame_dev),
                    c.pressure_iters, fl.ptr(self.rad), fl.ptr(self.delta_r),
                    fl.ptr(self.face_ids), fl.ptr(self.fsc), c.pressure_k, c.pressure_exp,
                    c.pressure_cap, c.pressure_max_nei, st)
        if c.brownian:
            fl.call("fos_batch_brownian", h, fl.ptr(self.pos), c.brownian_strength, float(c.L), 0,
                    fl.ptr(self.frame_dev), c.brownian_mode, st)
        for p in range(c.pbd_max):   # foams past their own pass count sit out
            self.rebuild_grid(self.ctl[0], p)
            fl.call("fos_batch_pbd_pass", h, fl.ptr(self.pos), fl.ptr(self.pos_tmp),
                    fl.ptr(self.rad), float(c.L), c.gap_factor(), c.max_disp_frac, c.disp_mult,
                    fl.ptr(self.cell_start), fl.ptr(self.cell_count), fl.ptr(self.cell_indices),
                    fl.ptr(self.cell_pr), fl.ptr(self.ctl[0]), p, st)
            self.pos, self.pos_tmp = self.pos_tmp, self.pos
        self.rebuild_grid()
        if ran:
            self.run_jfa(graph=True)
            fl.call("fos_batch_count_zero", h, fl.ptr(self.fsc), fl.ptr(self.zeros), st)
        fl.call("fos_batch_control_post", h, fl.ptr(self.zeros), fl.ptr(self.frame_dev),
                c.warmstart_frames, float(c.fsc_zero_rate), c.adjustment_frames, int(ran),
                fl.ptr(self.settle_dev), fl.ptr(self.ctl[2]), st)
        fl.call("fos_batch_fsc_targets", h, fl.ptr(self.rad), fl.ptr(self.rad_target),
                fl.ptr(self.fsc), fl.ptr(self.fsc_ema), c.growth_pct, c.deadband,
                fl.ptr(self.ctl[2]), st)
        fl.call("fos_batch_frame_advance", fl.ptr(self.frame_dev), st)

    def capture(self):
        """Records the frame as hip graphs (one with the JFA and, when the
        cadence skips frames, one without); replay() then runs one frame per
        graph launch.  Run eager step()s first: they size the face-row tables
        the graphs write.  Eager steps after a capture need a new capture."""
        c = self.cfg
        if c.pbd_max % 2:
            raise NotImplementedError("a captured frame needs an even PBD pass bound")
        dev = self.pos.device
        self.frame_dev = torch.tensor([self.frame], dtype=torch.int32, device=dev)
        self.settle_dev = torch.tensor(self.settle.astype(np.int32), device=dev)
        self.rmax_d = torch.tensor([float(cb.r_max) for cb in self.cfgs], dtype=torch.float64,
                                   device=dev)
        span = c.jfa_warm + max(1, c.jfa_every) + 1
        kinds = sorted({c.jfa_runs(f) for f in range(self.frame, self.frame + span)})
        torch.cuda.synchronize()
        self.graphs = {}
        for ran in kinds:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._frame_device(ran)
            self.graphs[ran] = g
        torch.cuda.synchronize()

    def replay(self, check=True):
        """One captured frame (enqueued; no host round trip unless check)."""
        ran = self.cfg.jfa_runs(self.frame)
        self.graphs[ran].replay()
        self.frame += 1
        if check:
            self.check()

    def check(self):
        """Fails loudly if any captured frame so far overflowed a face-row
        table or ended finalize before its fixed point (sticky device flags;
        one host round trip)."""
        out = (ctypes.c_int32 * 2)()
        fl.call("fos_batch_flags", self.ctx.handle, out, fl.stream())
        if out[0] or out[1]:
            raise RuntimeError(f"captured frames up to {self.frame - 1}: face-row table overflow "
                               f"({out[0]}) or finalize short of its fixed point after "
                               f"{FIXED_PAIRS} round pairs ({out[1]}); step() eagerly and "
                               "capture again")

    def snapshot(self, b):
        """Foam b's state with foam-local particle ids (foam.FoamSim.snapshot
        keys)."""
        torch.cuda.synchronize()
        p0 = int(self.p0[b])
        out = {}
        for k in ("pos", "rad", "rad_target", "lmd", "face_ids", "face_counts", "fsc", "fsc_ema",
                  "neighbor_list", "neighbor_count", "used_count", "prev_scores"):
            a = self.foam(b, getattr(self, k)).cpu().numpy().copy()
            if k in ("face_ids", "neighbor_list"):
                a[a >= 0] -= p0
            out[k] = a
        out["frame"] = np.int64(self.frame)
        settle = self.settle_dev.cpu().numpy() if getattr(self, "graphs", None) else self.settle
        out["settle"] = np.int64(settle[b])
        return out
