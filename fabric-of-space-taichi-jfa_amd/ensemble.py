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
        self.ctx = fl.context(n, int(self.res_of.max()), int(self.c0[-1]))
        self.frame = 0

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                fl.call("fos_batch_destroy", h)
            except Exception:
                pass
            self.handle = None

    def foam(self, b, t):
        """Foam b's rows of a batch tensor indexed by particle."""
        return t[int(self.p0[b]):int(self.p0[b + 1])]

    def labels(self, b, t=None):
        """Foam b's label block (res^3, batch-global ids) of grid or t."""
        t = self.grid if t is None else t
        return t[int(self.v0[b]):int(self.v0[b + 1])]

    def run_jfa(self):
        """jfa.run_jfa for every foam (jfa.py:1078-1132), batched."""
        c, ctx, st, h = self.cfg, self.ctx.handle, fl.stream(), self.handle
        fl.call("fos_batch_jfa_init", h, ctx, fl.ptr(self.pos), fl.ptr(self.rad), float(c.L),
                float(c.beta), fl.ptr(self.grid), None, st)
        fl.call("fos_batch_jfa_passes", h, ctx, float(c.L), int(c.early_exit), c.th_pass_delta,
                c.min_passes, fl.ptr(self.grid), fl.ptr(self.temp), fl.ptr(self.info), st)
        fl.call("fos_batch_majority", h, fl.ptr(self.grid), fl.ptr(self.filtered), st)
        fl.call("fos_batch_face_accumulate", h, ctx, fl.ptr(self.filtered),
                fl.ptr(self.face_ids), fl.ptr(self.face_counts), st)
        fl.call("fos_batch_finalize_faces", h, ctx, fl.ptr(self.rad), fl.ptr(self.face_ids),
                fl.ptr(self.face_counts), fl.ptr(self.prev_scores), fl.ptr(self.neighbor_list),
                fl.ptr(self.neighbor_count), fl.ptr(self.used_count), fl.ptr(self.fsc), st)
        fl.call("fos_fsc_ema", self.n, c.ema_alpha, fl.ptr(self.fsc), fl.ptr(self.fsc_ema), st)
