"""Constants of the per-frame foam step and the workload presets.

Module-level names mirror the hot-path subset of the reference config.py
(same names, same default values) so the drop-in modules read them the way
the reference kernels do.  Unlike the reference, where Taichi bakes them in at
JIT time, they are read at call time: ``apply(cfg)`` rebinds them for a run.

``FoamConfig`` is the headless driver's configuration; ``preset(name)`` gives
the benchmark workloads:
  c1   N=2,000, 64^3, uniform unit box, 20 frames, JFA every frame
  c2   N=10,000, 128^3, 200 frames, JFA every frame, Gaussian Brownian noise
       and the OU jitter (headline)
  c2c  as c2 with JFA on a 1/5 cadence after a 20-frame warm start
  c2t  as c2 with the reference's temporal-skip JFA scheduler
  c3   N=200,000 clustered, 384^3, 50 frames
  c5   N=50,000, 256^3, radii U[0.005, 0.011] (learned face predictor workload)
  c5s  as c5 with the predictor between JFA judge frames (every 8th)
"""
import dataclasses
import math

# ---- reference config.py values (config.py line numbers) ------------------
DOMAIN_SIZE = 0.189            # config.py:21
R_MIN = 0.002                  # config.py:49
R_MAX = 0.010                  # config.py:50
PBC_ENABLED = True             # config.py:57
JFA_EMA_ALPHA = 0.33           # config.py:81
FSC_LOW, FSC_HIGH = 8, 20      # config.py:90-91
GROWTH_PCT = 0.10              # config.py:92
ADJUSTMENT_FRAMES = 15         # config.py:93
MAX_STEP_PCT = 0.12            # config.py:96
FSC_DEADBAND = 1.0             # config.py:100
BACKPRESSURE_MODE = "local"    # config.py:101
PRESSURE_K = 0.10              # config.py:112
PRESSURE_EXP = 3.0             # config.py:113
PRESSURE_PAIR_CAP = 0.02       # config.py:114
MAX_EQ_NEI = 10                # config.py:115
BROWNIAN_STRENGTH = 0.0002     # config.py:123
WARMSTART_FRAMES = 30          # config.py:129
FSC_ZERO_RATE_THRESH = 0.10    # config.py:130
POWER_BETA = 1.0               # config.py:143
MIN_FACE_VOXELS = 2            # config.py:151
JFA_RES_MIN = 192              # config.py:156
JFA_RES_MAX = 320              # config.py:157
JFA_VOXEL_SCALE = 2.8          # config.py:158
JFA_ADAPTIVE_ENABLED = True    # config.py:163
JFA_VOXELS_PER_DIAMETER = 12.0 # config.py:164
JFA_CADENCE = 5                # config.py:172
JFA_WARMSTART_FRAMES = 30      # config.py:173
JFA_TILE_SIZE = 16             # config.py:182
JFA_DIRTY_HALO = 1             # config.py:183
JFA_DIRTY_POS_THRESHOLD = 0.5  # config.py:189
JFA_DIRTY_RAD_THRESHOLD = 0.25 # config.py:190
JFA_TEMPORAL_SKIP_ENABLED = True  # config.py:200
TH_RSIGMA = 0.004              # config.py:201
TH_MARGIN_SAFE = 0.4           # config.py:202
TH_FSC_DRIFT = 0.05            # config.py:203
MAX_SKIP_FRAMES = 15           # config.py:204
SKIP_BAND_CHANGE_LOCKOUT = 2   # config.py:205
OUT_OF_BAND_FRAC = 0.10        # run.py:976 (the out-of-band guard)
JFA_EARLY_EXIT_ENABLED = True  # config.py:208
TH_PASS_DELTA = 0.006          # config.py:209
MIN_JFA_PASSES = 3             # config.py:210
CONTACT_TOL = 0.035            # config.py:254
R_TYPICAL = 0.015              # config.py:289
CELL_SIZE = 2.0 * (1.0 + CONTACT_TOL) * R_TYPICAL   # config.py:290-291
GRID_RES = max(3, int(math.ceil(DOMAIN_SIZE / CELL_SIZE)))  # config.py:317
EPS = 1e-8                     # config.py:324
PBD_BASE_PASSES = 4            # config.py:342
PBD_MAX_PASSES = 8             # config.py:343
PBD_ADAPTIVE_SCALE = 20.0      # config.py:344
MAX_DISPLACEMENT_FRAC = 0.2    # config.py:349
DISPLACEMENT_MULTIPLIER = 2.0  # config.py:350
GAP_FRACTION = 0.015           # config.py:353
GLOBAL_DAMPING = 0.995         # config.py:384
DT = 0.016                     # config.py:400
XSPH_EPSILON = 0.03            # config.py:407
JITTER_ENABLED = True          # config.py:493
JITTER_RMS = 0.10              # config.py:494
JITTER_TAU = 1.0               # config.py:496
MAX_DRIFT_FRACTION = 0.20      # config.py:498
TOPO_EMA_ALPHA = 0.1           # config.py:455
TOPO_BATCHES = 12              # config.py:458
TOPO_PAIR_SUBSAMPLE_Q = 3      # config.py:459
TOPO_DEG_LOW = 10              # config.py:463
TOPO_DEG_HIGH = 18             # config.py:464
TOPO_MAX_RADIUS_MULTIPLE = 2.2 # config.py:467
MAX_TOPO_NEIGHBORS = 28        # config.py:468
TOPO_PAIR_SUBSAMPLING_STRIDE = 3     # config.py:471
TOPO_EARLY_EXIT_HIGH_DEGREE = True   # config.py:472
KNN_TOPO_K = 14                # config.py:487
HALF_L = 0.5 * DOMAIN_SIZE     # config.py:213
INV_L = 1.0 / DOMAIN_SIZE      # config.py:214
RNG_SEED = 0                   # Philox key for the noise kernels


def f32(x):
    """Rounds a Python float to fp32, as Taichi does for kernel constants."""
    import numpy as np
    return float(np.float32(x))


@dataclasses.dataclass
class FoamConfig:
    name: str = "custom"
    n: int = 2000
    L: float = 1.0
    res: int = 64
    num_passes: int = 0              # 0 -> ceil(log2 res) + 1 (jfa.py:224)
    early_exit: bool = False
    th_pass_delta: float = TH_PASS_DELTA
    min_passes: int = MIN_JFA_PASSES
    beta: float = POWER_BETA
    ema_alpha: float = JFA_EMA_ALPHA
    # radius clamps and the hash cell (2 * r_max)
    r_min: float = 0.01
    r_max: float = 0.04
    grid_res: int = 0                # 0 -> max(3, ceil(L / (2 r_max)))
    # PBD (run.py:712-714, 829-846)
    pbd_base: int = PBD_BASE_PASSES
    pbd_max: int = PBD_MAX_PASSES
    pbd_scale: float = PBD_ADAPTIVE_SCALE
    gap_fraction: float = GAP_FRACTION
    max_disp_frac: float = MAX_DISPLACEMENT_FRAC
    disp_mult: float = DISPLACEMENT_MULTIPLIER
    # FSC controller (run.py:733-751, 1184-1231)
    f_low: int = 10
    f_high: int = 16
    growth_pct: float = GROWTH_PCT
    adjustment_frames: int = ADJUSTMENT_FRAMES
    max_step_pct: float = MAX_STEP_PCT
    deadband: float = FSC_DEADBAND
    backpressure_mode: int = 2       # 0 off, 1 global, 2 local
    contact_tol: float = CONTACT_TOL
    warmstart_frames: int = WARMSTART_FRAMES
    fsc_zero_rate: float = FSC_ZERO_RATE_THRESH
    # pressure (run.py:755-771)
    pressure: bool = True
    pressure_k: float = PRESSURE_K
    pressure_exp: float = PRESSURE_EXP
    pressure_cap: float = PRESSURE_PAIR_CAP
    pressure_max_nei: int = MAX_EQ_NEI
    pressure_iters: int = 1
    # noise
    brownian: bool = False
    brownian_mode: int = 0           # 0 uniform (reference form), 1 Gaussian
    brownian_strength: float = BROWNIAN_STRENGTH
    ou_jitter: bool = False
    seed: int = RNG_SEED
    # which frames run the JFA: "cadence" (every `jfa_every` frames once
    # frame >= jfa_warm) or "temporal" (the reference's temporal skip,
    # run.py:902-1018: warm start, watchdog, band-change lockout, activity
    # and out-of-band triggers)
    jfa_schedule: str = "cadence"
    jfa_every: int = 1
    jfa_warm: int = 0
    skip_warmstart: int = JFA_WARMSTART_FRAMES
    max_skip: int = MAX_SKIP_FRAMES
    th_rsigma: float = TH_RSIGMA
    th_margin: float = TH_MARGIN_SAFE
    th_fsc_drift: float = TH_FSC_DRIFT
    oob_frac: float = OUT_OF_BAND_FRAC
    band_lockout: int = SKIP_BAND_CHANGE_LOCKOUT
    # adaptive JFA resolution on JFA frames (run.py:1049-1068): round(L *
    # voxels_per_diameter / (2 r_mean)) clamped to [res_min, res_max]
    adaptive_res: bool = False
    voxels_per_diameter: float = JFA_VOXELS_PER_DIAMETER
    res_min: int = JFA_RES_MIN
    res_max: int = JFA_RES_MAX
    frames: int = 20
    # learned face predictor: JFA (the judge) every `judge_every` frames, the
    # predictor in between (0 = plain JFA); fall back to the JFA on every
    # frame while the last judged frame's FSC error (judge_on) exceeds
    # max_fsc_err
    judge_every: int = 0
    max_fsc_err: float = 0.02
    # what the judge bounds by max_fsc_err: "per_particle" = sum |FSC_pred -
    # FSC_jfa| / sum FSC_jfa (every particle's own count), "mean" = |sum
    # FSC_pred - sum FSC_jfa| / sum FSC_jfa (the frame's mean FSC, where
    # over- and under-counts cancel)
    judge_on: str = "per_particle"
    # candidate pairs of the predictor: |d| < pred_mult * (ri + rj)
    pred_mult: float = 1.6

    def passes(self):
        return self.num_passes or int(math.ceil(math.log2(self.res))) + 1

    def hash_res(self):
        if self.grid_res:
            return self.grid_res
        return max(3, int(math.ceil(self.L / (2.0 * self.r_max))))

    def voxel(self):
        return f32(self.L / self.res)

    def gap_factor(self):
        return f32(1.0 + self.gap_fraction)

    def alpha(self):
        # run.py:440-452 alpha_for_frames
        return f32(1.0 - 0.05 ** (1.0 / max(1, self.adjustment_frames)))

    def jfa_runs(self, frame):
        if frame < self.jfa_warm:
            return True
        return (frame - self.jfa_warm) % max(1, self.jfa_every) == 0

    def max_jfa_res(self):
        """The largest resolution a JFA frame may run at."""
        return max(self.res, self.res_max) if self.adaptive_res else self.res


def preset(name):
    """Benchmark workloads (uniform unit box unless stated)."""
    if name == "c1":
        return FoamConfig(name="c1", n=2000, res=64, r_min=0.01, r_max=0.04,
                          pbd_base=8, pbd_max=8, f_low=10, f_high=16,
                          pressure_iters=10, frames=20, warmstart_frames=5)
    if name in ("c2", "c2c", "c2t"):
        # the reference frame also runs the OU jitter every frame (run.py:857-867)
        cfg = FoamConfig(name=name, n=10000, res=128, r_min=0.006, r_max=0.024,
                         pbd_base=PBD_BASE_PASSES, pbd_max=12, f_low=12, f_high=16,
                         pressure_iters=20, brownian=True, brownian_mode=1,
                         brownian_strength=1e-4, ou_jitter=True, frames=200)
        if name == "c2c":
            cfg.jfa_every, cfg.jfa_warm = 5, 20
        if name == "c2t":
            cfg.jfa_schedule = "temporal"
        return cfg
    if name == "c3":
        return FoamConfig(name="c3", n=200000, res=384, r_min=0.0015, r_max=0.015,
                          pbd_base=PBD_BASE_PASSES, pbd_max=PBD_MAX_PASSES,
                          f_low=12, f_high=16, pressure_iters=1, frames=50)
    if name in ("c5", "c5s"):
        # the learned face predictor's workload: N=50,000, radii U[0.005, 0.011],
        # JFA 256^3; c5s decimates the JFA to every 8th frame (the judge)
        cfg = FoamConfig(name=name, n=50000, res=256, r_min=0.004, r_max=0.011,
                         pbd_base=PBD_BASE_PASSES, pbd_max=12, f_low=12, f_high=16,
                         pressure_iters=20, brownian=True, brownian_mode=1,
                         brownian_strength=1e-4, frames=50)
        # Radii U[0.005, 0.011] fill ~14% of the box at N=50,000, so faces
        # reach well beyond 1.6 (ri + rj); 3.5 gives the ~40 candidates per
        # particle the workload is sized for.
        cfg.pred_mult = 3.5
        if name == "c5s":
            cfg.judge_every = 8
        return cfg
    raise KeyError(name)


def apply(cfg):
    """Rebinds the module constants the drop-in modules read to cfg."""
    g = globals()
    g["DOMAIN_SIZE"] = cfg.L
    g["HALF_L"] = 0.5 * cfg.L
    g["INV_L"] = 1.0 / cfg.L
    g["R_MIN"], g["R_MAX"] = cfg.r_min, cfg.r_max
    g["GRID_RES"] = cfg.hash_res()
    g["CELL_SIZE"] = cfg.L / cfg.hash_res()
    g["JFA_RES_MIN"] = min(JFA_RES_MIN, cfg.res)
    g["JFA_RES_MAX"] = max(JFA_RES_MAX, cfg.res)
    g["JFA_EARLY_EXIT_ENABLED"] = cfg.early_exit
    g["RNG_SEED"] = cfg.seed
