"""Headless per-frame foam step on one GPU (the caller contract of run.py).

One ``FoamSim.step()`` reproduces the order of the reference main loop
(run.py:700-1261) without GUI, render or prints:

  0  rebuild_grid                                   run.py:704
  1  compute_max_overlap -> adaptive PBD passes     run.py:709-714
  2  settle: nudge_radii_adaptive_ema               run.py:733-751
  3  equilibrate_pressure x pressure_iters          run.py:755-771
  4  apply_brownian_motion                          run.py:807-813
  5  passes x (rebuild_grid, project_overlaps)      run.py:829-846
  6  OU jitter (apply_brownian + integrate_jitter)  run.py:857-867
  7  rebuild_grid                                   run.py:874
  8  JFA (on cadence frames) -> FSC, EMA            run.py:1074
  9  warm-start gate, set_fsc_targets, settle reset run.py:1184-1231

XSPH and global damping act on a velocity field that the reference never
makes non-zero (rescue is off), so they are not run here (the drop-in
``dynamics`` module still exposes them).  PBD is the Jacobi form: every
particle reads the pass-start positions.

All work is enqueued on torch's current stream through libfos.  The host
reads the device once per frame for the overlap maximum (it sets the PBD pass
count, as in the reference) and once per JFA frame where the face rows are
finalized (row-table overflow and the fixed point).  The warm-start gate and
the settle counter it resets live on the device, so steps 2-4 (which need
neither the overlap maximum nor the gate on the host) are queued before the
host waits for the overlap maximum, behind an event rather than a drain of
the stream: the GPU works through them while the host reads.
"""
import ctypes
import math

import numpy as np
import torch

import fos_lib as fl
import foam_config as fc

MAXN = 32


class FrameInfo(dict):
    """What step() returns: frame, pbd_passes, max_depth, jfa, growth (and the
    JFA telemetry when on).  "growth" is decided on the device; reading it
    before the next frame has read it back costs one device read."""

    def __init__(self, sim, **kw):
        super().__init__(**kw)
        self._sim = sim

    def __getitem__(self, k):
        if k == "growth" and dict.get(self, "growth") is None:
            self._sim.resolve()
        return dict.__getitem__(self, k)

    def get(self, k, default=None):
        return self[k] if k in self else default


class FoamSim:
    def __init__(self, cfg, pos, rad, device="cuda", telemetry=False):
        fl.require_gpu()
        self.cfg = cfg
        self.telemetry = telemetry
        n = int(cfg.n)
        dev = torch.device(device)
        self.n = n
        p0 = torch.as_tensor(np.asarray(pos, np.float32).reshape(n, 3), device=dev).contiguous()
        # positions ping-pong between two buffers (PBD reads one, writes the
        # other); _cur says which holds the positions
        self._pair = [p0, torch.empty_like(p0)]
        self._cur = 0
        self.rad = torch.as_tensor(np.asarray(rad, np.float32).reshape(n), device=dev).contiguous()
        self.rad_target = self.rad.clone()
        self.lmd = torch.zeros(n, dtype=torch.float32, device=dev)
        self.delta_r = torch.zeros(n, dtype=torch.float32, device=dev)
        self.v_jit = torch.zeros(n, 3, dtype=torch.float32, device=dev)
        self.G = cfg.hash_res()
        C = self.G ** 3
        self.cell_count = torch.zeros(C, dtype=torch.int32, device=dev)
        self.cell_start = torch.zeros(C, dtype=torch.int32, device=dev)
        self.cell_indices = torch.zeros(max(n, 1), dtype=torch.int32, device=dev)
        # slot records (position, radius per CSR slot) for the PBD passes
        self.cell_pr = torch.zeros(max(n, 1), 4, dtype=torch.float32, device=dev)
        self.slot_records = True
        self.res = int(cfg.res)
        if cfg.jfa_schedule not in ("cadence", "temporal"):
            raise ValueError(f"jfa_schedule must be 'cadence' or 'temporal', not {cfg.jfa_schedule!r}")
        if cfg.jfa_schedule == "temporal" and cfg.judge_every > 0:
            raise ValueError("the temporal JFA schedule and the predictor's judge schedule "
                             "are exclusive")
        # the label buffers hold the largest grid a JFA frame may use
        # (adaptive resolution: res_max)
        self.max_res = int(cfg.max_jfa_res())
        V = self.max_res ** 3
        self.grid = torch.full((V,), -1, dtype=torch.int32, device=dev)
        self.temp = torch.full((V,), -1, dtype=torch.int32, device=dev)
        self.filtered = torch.full((V,), -1, dtype=torch.int32, device=dev)
        self.face_ids = torch.full((n, MAXN), -1, dtype=torch.int32, device=dev)
        self.face_counts = torch.zeros(n, MAXN, dtype=torch.int32, device=dev)
        self.fsc = torch.zeros(n, dtype=torch.int32, device=dev)
        self.fsc_ema = torch.zeros(n, dtype=torch.float32, device=dev)
        self.neighbor_list = torch.full((n, MAXN), -1, dtype=torch.int32, device=dev)
        self.neighbor_count = torch.zeros(n, dtype=torch.int32, device=dev)
        self.used_count = torch.zeros(n, dtype=torch.int32, device=dev)
        self.prev_scores = torch.zeros(n, MAXN, dtype=torch.float32, device=dev)
        self.info = torch.zeros(2 + 64, dtype=torch.int32, device=dev)
        self.pstats = torch.zeros(2, dtype=torch.float32, device=dev)
        # [0] overlap maximum of the frame, [1] the warm-start gate (1.0 =
        # growth) of the last frame that decided it on the device
        self.maxd = torch.zeros(2, dtype=torch.float32, device=dev)
        self.ctx = fl.context(n, self.max_res, C)
        # scheduler (temporal skip, adaptive resolution): positions and
        # radii at the last decision, and the statistics' host copy
        self.pos_prev = torch.zeros(n, 3, dtype=torch.float32, device=dev)
        self.rad_prev = torch.zeros(n, dtype=torch.float32, device=dev)
        self.sched_host = torch.zeros(8, dtype=torch.float64).pin_memory()
        self.frame = 0
        # settle counter (run.py:744-751, 1231) on the device, ping-pong:
        # slot sidx enters the next frame.  settle_force: the last frame set
        # it to adjustment_frames on the host (no gate); settle_gate: the
        # last frame's device gate still has to be applied to it.
        self.settle_dev = torch.zeros(2, dtype=torch.int32, device=dev)
        self.sidx = 0
        self.settle_force = False
        self.settle_gate = False
        # the overlap maximum and gate, copied to the host without a drain
        self.maxd_host = torch.zeros(2, dtype=torch.float32).pin_memory()
        self.growth = True
        self.pending = None        # FrameInfo whose gate is still on the device
        # (positions tensor, its version, hash grid edge) when the CSR was
        # last built at a frame's end; the next frame reuses that CSR while
        # nothing has written the positions since
        self.grid_key = None
        self.last = {}
        # optional list of ctypes arrays of passes+1 hipEvent_t handles (or
        # None); each JFA frame consumes one entry and, when it is an array,
        # times its pass kernels on the stream
        self.pass_events = None
        # optional callback(sim) after every JFA (training-data collection)
        self.on_jfa = None
        # learned face predictor (cfg.judge_every > 0): set_predictor()
        self.predictor = None
        self.fallback = False
        self.judged = []          # (frame, fsc error, mean-FSC error) per judge frame
        self.pred_fsc = torch.zeros(n, dtype=torch.int32, device=dev)
        # the plain-JFA frame runs as one library call (fos_frame_step) unless
        # native is False or a predictor or an on_jfa hook is set
        self.native = True
        self._fdesc = self._fbufs = None
        self._fstate = fl.FrameState()
        self._sched_since = 0
        self.jfa_res = self.res

    @property
    def pos(self):
        return self._pair[self._cur]

    @property
    def pos_tmp(self):
        return self._pair[1 - self._cur]

    def __del__(self):
        st = getattr(self, "_fstate", None)
        if st is not None and st.event and fl._lib is not None:
            try:
                fl._lib.fos_frame_release(ctypes.byref(st))
            except Exception:
                pass

    def set_predictor(self, params, mult=None):
        """Enables the predictor path with fp32 MLP parameters (numpy) and the
        candidate radius multiple they were trained with."""
        from face_predictor import FacePredictor
        c = self.cfg
        self.predictor = FacePredictor(params, self.n, c.L, c.r_max, device=self.pos.device,
                                       mult=c.pred_mult if mult is None else mult)
        self.r_scale = 0.5 * (c.r_min + c.r_max)

    # -- building blocks ------------------------------------------------------
    def grid_token(self):
        """Identity of the positions the CSR describes: torch bumps a tensor's
        version on every in-place write it makes (copy_, indexing stores)."""
        return (self.pos.data_ptr(), self.pos._version, self.G)

    def rebuild_grid(self, pr=False):
        """grid.rebuild_grid; pr=True also writes the slot records the PBD
        pass reads (when slot_records is on)."""
        c = self.cfg
        if pr and self.slot_records:
            fl.call("fos_grid_build_pr", self.ctx.handle, fl.ptr(self.pos), fl.ptr(self.rad),
                    self.n, self.G, c.L, fl.ptr(self.cell_count), fl.ptr(self.cell_start),
                    fl.ptr(self.cell_indices), fl.ptr(self.cell_pr), fl.stream())
            return
        fl.call("fos_grid_build", self.ctx.handle, fl.ptr(self.pos), self.n, self.G, c.L,
                fl.ptr(self.cell_count), fl.ptr(self.cell_start), fl.ptr(self.cell_indices),
                fl.stream())

    def max_overlap(self):
        c = self.cfg
        fl.call("fos_max_overlap", fl.ptr(self.pos), fl.ptr(self.rad), self.n, self.G, c.L,
                fl.ptr(self.cell_start), fl.ptr(self.cell_count), fl.ptr(self.cell_indices),
                fl.ptr(self.lmd), fl.ptr(self.maxd), fl.stream())

    def pbd_pass(self, cells=False):
        """One PBD pass; cells=True also runs the wrap-and-hash step of the
        grid rebuild that follows it (finish with rebuild_from_cells)."""
        c = self.cfg
        args = (fl.ptr(self.pos), fl.ptr(self.pos_tmp), fl.ptr(self.rad), self.n, self.G, c.L,
                c.gap_factor(), c.max_disp_frac, c.disp_mult, fl.ptr(self.cell_start),
                fl.ptr(self.cell_count), fl.ptr(self.cell_indices), fl.stream())
        if cells:
            pr = fl.ptr(self.cell_pr) if self.slot_records else None
            fl.call("fos_pbd_pass_cells", self.ctx.handle, *args[:-1], pr, args[-1])
        else:
            fl.call("fos_pbd_pass", *args)
        self._cur = 1 - self._cur

    def rebuild_from_cells(self, pr=False):
        rec = pr and self.slot_records
        fl.call("fos_grid_build_from_cells", self.ctx.handle, fl.ptr(self.pos),
                fl.ptr(self.rad), self.n, self.G, fl.ptr(self.cell_count),
                fl.ptr(self.cell_start), fl.ptr(self.cell_indices),
                fl.ptr(self.cell_pr) if rec else None, fl.stream())

    def pressure(self, frame):
        c = self.cfg
        fl.call("fos_pressure_run", self.ctx.handle, self.n, frame, c.pressure_iters,
                fl.ptr(self.rad), fl.ptr(self.delta_r), fl.ptr(self.face_ids), fl.ptr(self.fsc),
                c.pressure_k, c.pressure_exp, c.pressure_cap, c.pressure_max_nei, c.r_min,
                c.r_max, None, fl.stream())

    def run_jfa(self, res=None):
        """jfa.run_jfa without the host copies (jfa.py:1078-1132), at the
        configured resolution or `res` (adaptive resolution)."""
        c, ctx, st = self.cfg, self.ctx.handle, fl.stream()
        res = self.res if res is None else int(res)
        voxel = c.voxel() if res == c.res else fc.f32(c.L / res)
        passes = c.passes() if res == c.res else int(math.ceil(math.log2(res))) + 1
        self.jfa_res = res
        fl.call("fos_jfa_init", ctx, fl.ptr(self.pos), fl.ptr(self.rad), self.n, res, c.L,
                voxel, c.beta, fl.ptr(self.grid), None, st)
        ev = self.pass_events.pop(0) if self.pass_events else None
        fl.call("fos_jfa_passes_timed", ctx, res, c.L, voxel, passes,
                int(c.early_exit), c.th_pass_delta, c.min_passes, fl.ptr(self.grid),
                fl.ptr(self.temp), fl.ptr(self.info), ev, st)
        fl.call("fos_majority", res, fl.ptr(self.grid), fl.ptr(self.filtered), st)
        fl.call("fos_jfa_faces", ctx, res, self.n, fl.ptr(self.filtered),
                fl.ptr(self.rad), voxel, fl.ptr(self.face_ids), fl.ptr(self.face_counts),
                fl.ptr(self.prev_scores), fl.ptr(self.neighbor_list),
                fl.ptr(self.neighbor_count), fl.ptr(self.used_count), fl.ptr(self.fsc), st)
        fl.call("fos_fsc_ema", self.n, c.ema_alpha, fl.ptr(self.fsc), fl.ptr(self.fsc_ema), st)

    def resolve(self, gate=None):
        """Records the device-decided warm-start gate of the last frame (read
        now unless the caller read it) as that frame's growth; the settle
        reset it implies is applied on the device (fos_settle_nudge)."""
        if self.pending is None:
            return
        if gate is None:
            gate = float(self.maxd[1].item())
        g = gate > 0.5
        self.growth = g
        dict.__setitem__(self.pending, "growth", g)
        self.pending = None

    def gate_growth(self, f):
        """run.py:1184-1231 after a JFA (or predicted) frame: the gate and the
        gated FSC targets on the device, no host read."""
        c, st = self.cfg, fl.stream()
        gate = self.maxd.data_ptr() + 4
        fl.call("fos_fsc_gate", self.ctx.handle, fl.ptr(self.fsc), self.n, f,
                c.warmstart_frames, c.fsc_zero_rate, gate, st)
        fl.call("fos_fsc_targets_gated", fl.ptr(self.rad), fl.ptr(self.rad_target),
                fl.ptr(self.fsc), fl.ptr(self.fsc_ema), self.n, c.f_low, c.f_high,
                c.growth_pct, c.r_min, c.r_max, c.deadband, gate, st)

    # -- one frame ------------------------------------------------------------
    def _frame_desc(self):
        """fos_frame_desc from the configuration, with the values the Python
        frame passes (fp32 where it passes fp32, double where Python floats
        enter double arithmetic)."""
        c = self.cfg
        return fl.FrameDesc(
            n=self.n, res=self.res, G=self.G, jfa_passes=c.passes(), L=c.L, voxel=c.voxel(),
            beta=c.beta, ema_alpha=c.ema_alpha, early_exit=int(c.early_exit),
            min_passes=c.min_passes, th_pass_delta=c.th_pass_delta, pbd_base=c.pbd_base,
            pbd_max=c.pbd_max, pbd_scale=float(c.pbd_scale), r_max_d=float(c.r_max),
            gap_factor=c.gap_factor(), max_disp_frac=c.max_disp_frac, disp_mult=c.disp_mult,
            alpha=c.alpha(), max_step_pct=c.max_step_pct, r_min=c.r_min, r_max=c.r_max,
            contact_tol=c.contact_tol, contact_tol_d=float(c.contact_tol),
            backpressure_mode=c.backpressure_mode, adjustment_frames=c.adjustment_frames,
            warmstart_frames=c.warmstart_frames, fsc_zero_rate=c.fsc_zero_rate, f_low=c.f_low,
            f_high=c.f_high, growth_pct=c.growth_pct, deadband=c.deadband,
            pressure=int(c.pressure), pressure_iters=c.pressure_iters,
            pressure_max_nei=c.pressure_max_nei, pressure_k=c.pressure_k,
            pressure_exp=c.pressure_exp, pressure_cap=c.pressure_cap,
            brownian=int(c.brownian), brownian_mode=c.brownian_mode,
            brownian_strength=c.brownian_strength, seed=c.seed, ou_jitter=int(c.ou_jitter),
            dt=fc.DT, jitter_rms=fc.JITTER_RMS, jitter_tau=fc.JITTER_TAU,
            drift_cap_frac=fc.f32(fc.MAX_DRIFT_FRACTION * c.gap_fraction),
            jfa_every=c.jfa_every, jfa_warm=c.jfa_warm, slot_records=int(self.slot_records),
            schedule=1 if c.jfa_schedule == "temporal" else 0, skip_warmstart=c.skip_warmstart,
            max_skip=c.max_skip, th_rsigma=float(c.th_rsigma), th_margin=float(c.th_margin),
            th_fsc_drift=float(c.th_fsc_drift), oob_frac=float(c.oob_frac),
            adaptive_res=int(c.adaptive_res), res_min=c.res_min, res_max=c.res_max,
            voxels_per_diameter=float(c.voxels_per_diameter), L_d=float(c.L))

    def _frame_bufs(self):
        b = fl.FrameBufs()
        b.pos[0], b.pos[1] = self._pair[0].data_ptr(), self._pair[1].data_ptr()
        for k in ("rad", "rad_target", "lmd", "delta_r", "v_jit", "cell_count", "cell_start",
                  "cell_indices", "cell_pr", "grid", "temp", "filtered", "face_ids",
                  "face_counts", "fsc", "fsc_ema", "neighbor_list", "neighbor_count",
                  "used_count", "prev_scores", "info", "maxd", "maxd_host", "settle_dev",
                  "pos_prev", "rad_prev", "sched_host"):
            setattr(b, k, getattr(self, k).data_ptr())
        return b

    def _frame_objs(self):
        if self._fdesc is None:
            self._fdesc, self._fbufs = self._frame_desc(), self._frame_bufs()
        return self._fdesc, self._fbufs

    def schedule_jfa(self):
        """The JFA decision of this frame (fos_jfa_schedule: the cadence, or
        the temporal skip of run.py:902-1018) and the resolution a JFA frame
        runs at (run.py:1049-1068 when adaptive).  Returns (run, res)."""
        d, b = self._frame_objs()
        s = self._fstate
        s.frame = self.frame
        run, res = ctypes.c_int(), ctypes.c_int()
        since = s.since
        fl.call("fos_jfa_schedule", self.ctx.handle, ctypes.byref(d), ctypes.byref(b),
                ctypes.byref(s), fl.ptr(self.pos), ctypes.byref(run), ctypes.byref(res),
                fl.stream())
        self._sched_since = since + 1
        return bool(run.value), int(res.value)

    def schedule_record(self):
        """The last scheduling decision: reason, frames since the last JFA as
        the reference prints it, and (temporal skip) the signals."""
        s = self._fstate
        rec = {"jfa_reason": fl.JFA_REASONS.get(int(s.reason), "?")}
        if self.cfg.jfa_schedule == "temporal":
            rec["jfa_since"] = int(s.since) if int(s.reason) == 0 else self._sched_since
            rec["jfa_signals"] = {"sigma_dr_over_r": float(s.signals[0]),
                                  "vmax_dt": float(s.signals[1]),
                                  "delta_fsc_mu": float(s.signals[2]),
                                  "out_of_band": float(s.signals[3]),
                                  "r_mean": float(s.signals[4])}
        return rec

    def set_band(self, f_low, f_high):
        """Changes the FSC band between frames; with the temporal skip a change
        forces the JFA for the lockout (run.py:1451-1453)."""
        c = self.cfg
        if (int(f_low), int(f_high)) == (c.f_low, c.f_high):
            return
        c.f_low, c.f_high = int(f_low), int(f_high)
        if c.jfa_schedule == "temporal":
            self._fstate.lockout = int(c.band_lockout)
        self._fdesc = None

    def _step_native(self):
        """step() as one fos_frame_step call (the same calls in the same
        order, issued from C++)."""
        c, f = self.cfg, self.frame
        s = self._fstate
        self._frame_objs()
        since = s.since
        s.frame, s.pos_cur, s.sidx = f, self._cur, self.sidx
        s.settle_force, s.settle_gate = int(self.settle_force), int(self.settle_gate)
        s.need_rebuild = int(self.cfg.L != 1.0 or self.grid_key != self.grid_token())
        s.pending = int(self.pending is not None)
        ev = self.pass_events.pop(0) if self.pass_events else None
        fl.call("fos_frame_step", self.ctx.handle, ctypes.byref(self._fdesc),
                ctypes.byref(self._fbufs), ctypes.byref(s), ev, fl.stream())
        if s.gate_read:
            self.resolve(float(s.gate))
        self._cur, self.sidx = s.pos_cur, s.sidx
        self.settle_force, self.settle_gate = bool(s.settle_force), bool(s.settle_gate)
        self.grid_key = self.grid_token()
        ran = bool(s.jfa_ran)
        self._sched_since = since + 1
        if ran:
            self.jfa_res = int(s.jfa_res) or self.res
        self.last = FrameInfo(self, frame=f, pbd_passes=int(s.pbd_passes),
                              max_depth=float(s.max_depth), jfa=ran,
                              growth=None if ran else True)
        self.last.update(self.schedule_record())
        if ran:
            self.last["jfa_res"] = self.jfa_res
            self.pending = self.last
        else:
            self.growth = True
        if self.telemetry and ran:
            self.last.update(self.jfa_telemetry())
        self.frame += 1
        return self.last

    def run(self, frames):
        """`frames` step()s; returns the last frame's record.  On the native
        path (no predictor, hook or telemetry) one fos_frame_run call issues
        them all from C++ -- the same calls in the same order -- so the host
        queues the next frame while the GPU finishes the last; only the last
        frame's record is kept."""
        if (frames <= 1 or not self.native or self.predictor is not None
                or self.on_jfa is not None or self.telemetry):
            last = None
            for _ in range(frames):
                last = self.step()
            return last
        # the growth of the last gated frame, read now (the run reads the
        # following ones itself)
        self.resolve()
        s = self._fstate
        self._frame_objs()
        f0 = self.frame
        s.frame, s.pos_cur, s.sidx = f0, self._cur, self.sidx
        s.settle_force, s.settle_gate = int(self.settle_force), int(self.settle_gate)
        s.need_rebuild = int(self.cfg.L != 1.0 or self.grid_key != self.grid_token())
        s.pending = 0
        evs = [self.pass_events.pop(0) if self.pass_events else None for _ in range(frames)]
        arr = (ctypes.c_void_p * frames)(
            *[None if e is None else ctypes.cast(e, ctypes.c_void_p).value for e in evs])
        fl.call("fos_frame_run", self.ctx.handle, ctypes.byref(self._fdesc),
                ctypes.byref(self._fbufs), ctypes.byref(s), frames, arr, fl.stream())
        if s.gate_read:
            self.growth = float(s.gate) > 0.5
        self._cur, self.sidx = s.pos_cur, s.sidx
        self.settle_force, self.settle_gate = bool(s.settle_force), bool(s.settle_gate)
        self.grid_key = self.grid_token()
        ran = bool(s.jfa_ran)
        if ran:
            self.jfa_res = int(s.jfa_res) or self.res
        self.last = FrameInfo(self, frame=f0 + frames - 1, pbd_passes=int(s.pbd_passes),
                              max_depth=float(s.max_depth), jfa=ran,
                              growth=None if ran else True)
        self.last["jfa_reason"] = fl.JFA_REASONS.get(int(s.reason), "?")
        if ran:
            self.pending = self.last
        else:
            self.growth = True
        self.frame = f0 + frames
        return self.last

    def step(self):
        if self.native and self.predictor is None and self.on_jfa is None:
            return self._step_native()
        c = self.cfg
        f = self.frame
        st = fl.stream()
        # 0-1: the rebuild at the end of the last frame (step 7) built this
        # CSR from these same positions unless something has written them
        # since (a restore, a caller's store).  With L = 1 the wrap is
        # idempotent in fp32, so rebuilding again would reproduce positions
        # and CSR bit for bit.  Writes made through the C ABI do not bump the
        # tensor's version: set grid_key = None after one.
        if self.cfg.L != 1.0 or self.grid_key != self.grid_token():
            self.rebuild_grid()
        self.max_overlap()
        # the overlap maximum (and the last frame's gate) go to the host
        # behind an event; settle, pressure and noise need neither on the
        # host and are queued before the host waits, so the GPU keeps busy
        self.maxd_host.copy_(self.maxd, non_blocking=True)
        ready_ev = torch.cuda.Event()
        ready_ev.record()
        # 2 settle (counter and gate on the device)
        sin, sout = self.settle_dev[self.sidx:], self.settle_dev[1 - self.sidx:]
        fl.call("fos_settle_nudge", fl.ptr(self.rad), fl.ptr(self.rad_target), fl.ptr(self.lmd),
                self.n, c.alpha(), c.max_step_pct, c.r_min, c.r_max, c.backpressure_mode,
                c.contact_tol, float(c.contact_tol), fl.ptr(self.maxd), fl.ptr(sin), fl.ptr(sout),
                c.adjustment_frames, int(self.settle_force), int(self.settle_gate), st)
        self.sidx = 1 - self.sidx
        self.settle_force = self.settle_gate = False
        # 3 pressure
        if c.pressure:
            self.pressure(f)
        # 4 Brownian; with PBD passes to follow (pbd_base > 0), the noise
        # kernel also runs the wrap-and-hash step of the first pass's rebuild
        ready = False
        if c.brownian:
            if c.pbd_base > 0:
                fl.call("fos_brownian_cells", self.ctx.handle, fl.ptr(self.pos), self.n,
                        c.brownian_strength, c.L, c.seed, f, c.brownian_mode, self.G, st)
                ready = True
            else:
                fl.call("fos_brownian", fl.ptr(self.pos), self.n, c.brownian_strength, c.L,
                        c.seed, f, c.brownian_mode, st)
        ready_ev.synchronize()
        max_depth = float(self.maxd_host[0].item())
        self.resolve(float(self.maxd_host[1].item()))
        passes = max(c.pbd_base, min(c.pbd_max,
                                     int(c.pbd_base + c.pbd_scale * (max_depth / (0.2 * c.r_max)))))
        # 5 PBD; a pass followed directly by a rebuild runs that rebuild's
        # wrap-and-hash step itself
        for p in range(passes):
            if ready:
                self.rebuild_from_cells(pr=True)
            else:
                self.rebuild_grid(pr=True)
            ready = p < passes - 1 or not c.ou_jitter
            self.pbd_pass(cells=ready)
        # 6 OU jitter (the mean radius of run.py:860-861 computed on the device)
        if c.ou_jitter:
            fl.call("fos_ou_jitter_mean", self.ctx.handle, fl.ptr(self.pos), fl.ptr(self.v_jit),
                    fl.ptr(self.rad), self.n, fc.DT, fc.JITTER_RMS, fc.JITTER_TAU,
                    fc.f32(fc.MAX_DRIFT_FRACTION * c.gap_fraction), c.L, c.seed, f, None, st)
        # 7
        if ready:
            self.rebuild_from_cells()
        else:
            self.rebuild_grid()
        self.grid_key = self.grid_token()
        # 8 JFA on cadence frames (or the predictor between judge frames)
        if self.predictor is not None and c.judge_every > 0:
            ran = self.fallback or f % c.judge_every == 0
            if ran:
                self.run_jfa()
                warm = self.predictor.warm
                self.predictor.predict(self.pos, self.rad, self.n, self.pred_fsc,
                                       r_mean=self.r_scale)
                self.predictor.remember_jfa(self.neighbor_list, self.neighbor_count, self.n)
                from face_predictor import fsc_error
                err, bias = fsc_error(self.pred_fsc, self.fsc)
                # the judge bounds the per-particle error (judge_on
                # "per_particle") or the frame's mean FSC ("mean"); both are
                # recorded.  Without a previous frame's rows (a cold start)
                # the hysteresis feature is empty, so that judgement says
                # nothing about the frames that follow (they start from this
                # JFA's rows): no fallback.
                judged = err if c.judge_on == "per_particle" else bias
                self.fallback = warm and judged > c.max_fsc_err
                self.judged.append((f, err, bias))
            else:
                self.predictor.predict(self.pos, self.rad, self.n, self.fsc, self.face_ids,
                                       self.face_counts, r_mean=self.r_scale)
                fl.call("fos_fsc_ema", self.n, c.ema_alpha, fl.ptr(self.fsc),
                        fl.ptr(self.fsc_ema), st)
            gated = True
        else:
            ran, res = self.schedule_jfa()
            if ran:
                self.run_jfa(res)
                if self.on_jfa is not None:
                    self.on_jfa(self)
            gated = ran
        # 9 controller: after a JFA the gate is decided on the device
        self.last = FrameInfo(self, frame=f, pbd_passes=passes, max_depth=max_depth, jfa=ran,
                              growth=None if gated else True)
        if self.predictor is None or c.judge_every <= 0:
            self.last.update(self.schedule_record())
            if ran:
                self.last["jfa_res"] = self.jfa_res
        if gated:
            self.gate_growth(f)
            self.pending = self.last
            self.settle_gate = True
        else:
            self.growth = True
            fl.call("fos_fsc_targets", fl.ptr(self.rad), fl.ptr(self.rad_target), fl.ptr(self.fsc),
                    fl.ptr(self.fsc_ema), self.n, c.f_low, c.f_high, c.growth_pct, c.r_min,
                    c.r_max, c.deadband, st)
            self.settle_force = True
        if self.telemetry and ran:
            self.last.update(self.jfa_telemetry())
        self.frame += 1
        return self.last

    def jfa_telemetry(self):
        """jfa.py:1139-1195 stats and jfa.py:1400-1426 validation."""
        c = self.cfg
        oi = (fl.I64 * 5)()
        of = (fl.D * 4)()
        fl.call("fos_jfa_stats", self.ctx.handle, self.n, fl.ptr(self.rad),
                c.voxel() if self.jfa_res == c.res else fc.f32(c.L / self.jfa_res),
                fl.ptr(self.fsc), fl.ptr(self.fsc_ema), fl.ptr(self.used_count),
                fl.ptr(self.face_ids), fl.ptr(self.face_counts), oi, of, fl.stream())
        vo = (fl.I64 * 3)()
        fl.call("fos_validate", self.ctx.handle, self.n, fl.ptr(self.neighbor_list),
                fl.ptr(self.neighbor_count), vo, fl.stream())
        info = self.info.cpu().numpy()
        passes = int(info[0])
        V = float(self.jfa_res) ** 3
        return {"num_passes": passes, "early_exit": bool(info[1]),
                "delta_history": [float(c) / V for c in info[2:2 + passes]],
                "p_dir": int(oi[0]), "p_und_accepted": int(oi[0]) // 2,
                "overflow_count": int(oi[1]),
                "overflow_pct": 100.0 * int(oi[1]) / max(1, self.n), "min_fsc": int(oi[2]),
                "max_fsc": int(oi[3]), "mean_fsc": of[0], "mean_fsc_ema": of[1],
                "mean_score": of[2], "std_score": of[3], "asymmetric_pairs": int(vo[0]),
                "validation_overflow": int(vo[1]), "self_loops": int(vo[2])}

    def log_lines(self):
        """Reference-format telemetry of the frame just stepped: the JFA
        summary (run.py:1095-1149) when the JFA ran with telemetry on, and
        the frame summary (run.py:1303-1318) every 100th frame."""
        import telemetry as tm
        c, last, out = self.cfg, self.last, []
        if c.jfa_schedule == "temporal" and "jfa_signals" in last:
            out += tm.jfa_schedule_lines(last, int(self._fstate.sched_frames), c.skip_warmstart,
                                         c.th_rsigma, c.th_margin, c.th_fsc_drift, c.max_skip)
        if self.telemetry and last.get("jfa") and "num_passes" in last:
            val = {"passed": last["validation_overflow"] == 0 and last["self_loops"] == 0,
                   "asymmetric_pairs": last["asymmetric_pairs"]}
            out += tm.jfa_step4_lines(last, val, 0.0, self.jfa_res, self.n)
        if self.frame % 100 == 0:
            out += tm.frame_lines(self.frame, self.n, last.get("pbd_passes", 0),
                                  last.get("max_depth", 0.0), self.fsc.cpu().numpy(),
                                  self.rad.cpu().numpy(), c.f_low, c.f_high)
        return out

    # -- state ------------------------------------------------------------------
    STATE = ("pos", "rad", "rad_target", "lmd", "v_jit", "face_ids", "face_counts", "fsc",
             "fsc_ema", "neighbor_list", "neighbor_count", "used_count", "prev_scores")

    def snapshot(self):
        self.resolve()
        torch.cuda.synchronize()
        out = {k: getattr(self, k).detach().cpu().numpy().copy() for k in self.STATE}
        out["frame"] = np.int64(self.frame)
        out["settle"] = np.int64(self.settle)
        s = self._fstate
        out["pos_prev"] = self.pos_prev.cpu().numpy().copy()
        out["rad_prev"] = self.rad_prev.cpu().numpy().copy()
        out["sched"] = np.array([s.sched_frames, s.since, s.have_prev, s.fsc_mu_set,
                                 s.fsc_mu_pending, s.lockout], np.int64)
        out["sched_fsc_mu"] = np.float64(s.last_fsc_mu)
        return out

    @property
    def settle(self):
        """The settle counter the next frame starts from (run.py:744-751,
        1231), read from the device."""
        if self.settle_force:
            return int(self.cfg.adjustment_frames)
        s = int(self.settle_dev[self.sidx].item())
        if self.settle_gate and float(self.maxd[1].item()) > 0.5:
            s = int(self.cfg.adjustment_frames)
        return s

    def restore(self, snap):
        for k in self.STATE:
            getattr(self, k).copy_(torch.from_numpy(np.ascontiguousarray(snap[k])))
        self.frame = int(snap["frame"])
        self.settle_dev[self.sidx] = int(snap["settle"])
        self.settle_force = self.settle_gate = False
        self.pending = None
        self.grid_key = None
        s = self._fstate
        if "sched" in snap:
            self.pos_prev.copy_(torch.from_numpy(np.ascontiguousarray(snap["pos_prev"])))
            self.rad_prev.copy_(torch.from_numpy(np.ascontiguousarray(snap["rad_prev"])))
            (s.sched_frames, s.since, s.have_prev, s.fsc_mu_set, s.fsc_mu_pending,
             s.lockout) = [int(v) for v in snap["sched"]]
            s.last_fsc_mu = float(snap["sched_fsc_mu"])
        else:
            s.sched_frames = s.since = s.have_prev = s.fsc_mu_set = s.fsc_mu_pending = 0
            s.lockout = 0
            s.last_fsc_mu = 0.0

    def save(self, path):
        np.savez(path, **self.snapshot())

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            self.restore({k: z[k] for k in z.files})

    def export_csv(self, path):
        """run.py:608-629 particle dump: ID,X,Y,Z,Radius,FSC."""
        p = self.pos.cpu().numpy()
        r = self.rad.cpu().numpy()
        f = self.fsc.cpu().numpy()
        with open(path, "w") as fh:
            fh.write("ID,X,Y,Z,Radius,FSC\n")
            for i in range(self.n):
                fh.write(f"{i},{p[i,0]:.6f},{p[i,1]:.6f},{p[i,2]:.6f},{r[i]:.6f},{int(f[i])}\n")
