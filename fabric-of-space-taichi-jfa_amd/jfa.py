"""Drop-in for the reference jfa.py: power-diagram Jump Flood and FSC.

Same function names, parameters and return types as the reference module;
fields are torch device tensors wrapped in ``fos_lib.Field`` (``to_numpy()``,
``[None]`` for scalars).  Every computation runs in libfos.so (HIP, gfx950);
without it the calls raise.

Differences a caller can see:
* ``jfa_grid``/``jfa_temp`` hold int32 owner ids (-1 = unassigned) instead of
  the reference's (id as float, r^2) pairs; the r^2 channel was never read
  back (jfa.py:549,565 recompute it from ``rad``).
* Fields are sized by ``init_jfa(max_particles, max_res)``; the resolution
  clamp [JFA_RES_MIN, JFA_RES_MAX] and MAX_PARTICLES are parameters.
* Every per-frame kernel follows the serial semantics of its reference loop
  (seed collisions -> lowest power distance then lowest index; face rows in
  first-encounter order; finalize in ascending particle order), so results
  do not depend on thread timing.
"""
import math

import numpy as np
import torch

import foam_config as _C
import fos_lib as _fl
from fos_lib import Field

# jfa.py:53-74
EXPECTED_FACE_SCALE = 1.5
T_ADD = 0.15
T_DROP = 0.10
MAX_PARTICLES = 50000
MAX_NEIGHBORS = 32

# jfa.py:80-91 (dynamic configuration)
JFA_RES = 128
JFA_NUM_PASSES = 8
VOXEL_SIZE = _C.DOMAIN_SIZE / JFA_RES
power_beta_current = None

jfa_grid = None
jfa_temp = None
face_area = None
neighbor_list = None
neighbor_count = None
neighbor_used_count = None
fsc = None
fsc_ema = None
label_changes = None
label_filtered = None
face_counts = None
face_ids = None
neighbor_prev_scores = None
tile_dirty = None
pos_prev = None
rad_prev = None
tiles_per_axis = 0

_max_res = None
_flat = {}
_info = None
_device = None


def _dev():
    return _device or torch.device("cuda")


def _L():
    return float(_C.DOMAIN_SIZE)


def init_jfa(max_particles=None, max_res=None, device="cuda"):
    """jfa.py:120-204: allocates the module fields (zero-initialised like
    Taichi fields) for up to max_particles and max_res^3 voxels."""
    global jfa_grid, jfa_temp, face_area, neighbor_list, neighbor_count, neighbor_used_count
    global fsc, fsc_ema, label_changes, label_filtered, face_counts, face_ids
    global neighbor_prev_scores, power_beta_current, tile_dirty, pos_prev, rad_prev
    global tiles_per_axis, MAX_PARTICLES, _max_res, _flat, _info, _device
    _fl.require_gpu()
    _device = torch.device(device)
    if max_particles is not None:
        MAX_PARTICLES = int(max_particles)
    _max_res = int(max_res if max_res is not None else _C.JFA_RES_MAX)
    d, n, k = _device, MAX_PARTICLES, MAX_NEIGHBORS
    V = _max_res ** 3
    _flat = {
        "grid": torch.zeros(V, dtype=torch.int32, device=d),
        "temp": torch.zeros(V, dtype=torch.int32, device=d),
        "filt": torch.zeros(V, dtype=torch.int32, device=d),
    }
    face_area = Field(torch.zeros(n, k, dtype=torch.int32, device=d))
    neighbor_list = Field(torch.zeros(n, k, dtype=torch.int32, device=d))
    neighbor_count = Field(torch.zeros(n, dtype=torch.int32, device=d))
    neighbor_used_count = Field(torch.zeros(n, dtype=torch.int32, device=d))
    fsc = Field(torch.zeros(n, dtype=torch.int32, device=d))
    fsc_ema = Field(torch.zeros(n, dtype=torch.float32, device=d))
    label_changes = Field(torch.zeros((), dtype=torch.int32, device=d))
    face_counts = Field(torch.zeros(n, k, dtype=torch.int32, device=d))
    face_ids = Field(torch.zeros(n, k, dtype=torch.int32, device=d))
    neighbor_prev_scores = Field(torch.zeros(n, k, dtype=torch.float32, device=d))
    power_beta_current = Field(torch.tensor(_C.POWER_BETA, dtype=torch.float32))
    max_tiles = (_max_res + _C.JFA_TILE_SIZE - 1) // _C.JFA_TILE_SIZE
    tile_dirty = Field(torch.zeros(max_tiles ** 3, dtype=torch.uint8, device=d))
    pos_prev = Field(torch.zeros(n, 3, dtype=torch.float32, device=d))
    rad_prev = Field(torch.zeros(n, dtype=torch.float32, device=d))
    _info = torch.zeros(2 + 64, dtype=torch.int32, device=d)
    tiles_per_axis = 0
    _bind_views()


def _bind_views():
    global jfa_grid, jfa_temp, label_filtered
    r = JFA_RES
    if r > _max_res:                 # bound by set_resolution once it fits
        jfa_grid = jfa_temp = label_filtered = None
        return
    V = r ** 3
    jfa_grid = Field(_flat["grid"][:V].view(r, r, r))
    jfa_temp = Field(_flat["temp"][:V].view(r, r, r))
    label_filtered = Field(_flat["filt"][:V].view(r, r, r))


def set_resolution(new_res: int):
    """jfa.py:207-227."""
    global JFA_RES, VOXEL_SIZE, JFA_NUM_PASSES, tiles_per_axis
    JFA_RES = max(_C.JFA_RES_MIN, min(int(new_res), _C.JFA_RES_MAX))
    if _max_res is not None and JFA_RES > _max_res:
        raise ValueError(f"resolution {JFA_RES} exceeds init_jfa(max_res={_max_res})")
    VOXEL_SIZE = _L() / JFA_RES
    JFA_NUM_PASSES = int(np.ceil(np.log2(JFA_RES))) + 1
    tiles_per_axis = (JFA_RES + _C.JFA_TILE_SIZE - 1) // _C.JFA_TILE_SIZE
    if _flat:
        _bind_views()


def _need_init():
    if fsc is None:
        raise RuntimeError("jfa.init_jfa() must be called first")


def _pr(x):
    return _fl.ptr(_fl.tensor_of(x))


def _ctx(n):
    return _fl.context(max(1, n), JFA_RES, 27)


def _beta():
    return float(power_beta_current[None])


def _voxel():
    return _C.f32(VOXEL_SIZE)


# ---- kernels --------------------------------------------------------------

def jfa_init(pos, rad, active_n):
    """jfa.py:485-520: clear both grids, seed every particle's voxel."""
    _need_init()
    n = int(active_n)
    _fl.call("fos_jfa_init", _ctx(n).handle, _pr(pos), _pr(rad), n, JFA_RES, _L(), _voxel(),
             _beta(), _pr(_flat["grid"]), _pr(_flat["temp"]), _fl.stream())


def _step(pos, rad, step_size, src, dst):
    _need_init()
    n = _fl.tensor_of(rad).shape[0]
    ctx = _ctx(n)
    _fl.call("fos_jfa_seed_table", ctx.handle, _pr(pos), _pr(rad), n, _beta(), _fl.stream())
    label_changes.tensor.zero_()
    _fl.call("fos_jfa_step", ctx.handle, JFA_RES, _L(), _voxel(), int(step_size),
             _pr(_flat[src]), _pr(_flat[dst]), _pr(label_changes), _fl.stream())


def jfa_step_ping0(pos, rad, step_size):
    """jfa.py:524-572: jfa_grid -> jfa_temp."""
    _step(pos, rad, step_size, "grid", "temp")


def jfa_step_ping1(pos, rad, step_size):
    """jfa.py:576-623: jfa_temp -> jfa_grid."""
    _step(pos, rad, step_size, "temp", "grid")


def jfa_copy_temp_to_grid():
    """jfa.py:627-632."""
    _need_init()
    _fl.call("fos_jfa_copy", JFA_RES, _pr(_flat["temp"]), _pr(_flat["grid"]), _fl.stream())


def majority_filter_labels():
    """jfa.py:640-699."""
    _need_init()
    _fl.call("fos_majority", JFA_RES, _pr(_flat["grid"]), _pr(_flat["filt"]), _fl.stream())


def accumulate_face_area(active_n):
    """jfa.py:703-762."""
    _need_init()
    n = int(active_n)
    _fl.call("fos_face_accumulate", _ctx(n).handle, JFA_RES, n, _pr(_flat["filt"]),
             _pr(face_ids), _pr(face_counts), _fl.stream())


def update_fsc_ema(active_n, alpha):
    """jfa.py:766-785."""
    _need_init()
    _fl.call("fos_fsc_ema", int(active_n), float(alpha), _pr(fsc), _pr(fsc_ema), _fl.stream())


def finalize_faces_kernel(active_n, pos, rad, voxel_size):
    """jfa.py:789-857 (pos is unused, as in the reference)."""
    _need_init()
    n = int(active_n)
    _fl.call("fos_finalize_faces", _ctx(n).handle, n, _pr(rad), _C.f32(voxel_size),
             _pr(face_ids), _pr(face_counts), _pr(neighbor_prev_scores), _pr(neighbor_list),
             _pr(neighbor_count), _pr(neighbor_used_count), _pr(fsc), _fl.stream())


def extract_faces(active_n):
    """jfa.py:864-925 is the superseded path the reference never calls."""
    raise NotImplementedError("extract_faces is the reference's unused old path")


def finalize_fsc(active_n):
    """jfa.py:928-958 is the superseded path the reference never calls."""
    raise NotImplementedError("finalize_fsc is the reference's unused old path")


# ---- dirty tiles (instrumentation only, jfa.py:966-1049) ---------------------

def mark_dirty_tiles(pos, rad, n):
    _need_init()
    _fl.call("fos_dirty_mark", _pr(pos), _pr(rad), _pr(pos_prev), _pr(rad_prev), int(n),
             JFA_RES, _L(), _voxel(), _C.JFA_TILE_SIZE, _C.JFA_DIRTY_HALO,
             _C.JFA_DIRTY_POS_THRESHOLD, _C.JFA_DIRTY_RAD_THRESHOLD, _pr(_tile_view()),
             _fl.stream())


def _tile_view():
    t = max(1, tiles_per_axis)
    return tile_dirty.tensor[: t ** 3]


def count_dirty_tiles():
    _need_init()
    out = (_fl.I64 * 1)()
    _fl.call("fos_dirty_count", _ctx(1).handle, _pr(_tile_view()), max(1, tiles_per_axis), out,
             _fl.stream())
    return int(out[0])


def clear_dirty_tiles():
    _need_init()
    _tile_view().zero_()


def update_tile_cache(pos, rad, n):
    _need_init()
    _fl.call("fos_tile_cache_update", _pr(pos), _pr(rad), int(n), _pr(pos_prev), _pr(rad_prev),
             _fl.stream())


# ---- main pipeline ------------------------------------------------------------

def run_jfa(pos, rad, active_n):
    """jfa.py:1056-1197: returns (fsc[:n] as np.int32, stats dict)."""
    _need_init()
    n = int(active_n)
    ctx = _ctx(n)
    st = _fl.stream()
    voxel = _voxel()
    early = bool(_C.JFA_EARLY_EXIT_ENABLED)
    _fl.call("fos_jfa_init", ctx.handle, _pr(pos), _pr(rad), n, JFA_RES, _L(), voxel, _beta(),
             _pr(_flat["grid"]), None, st)
    _fl.call("fos_jfa_passes", ctx.handle, JFA_RES, _L(), voxel, JFA_NUM_PASSES, int(early),
             float(_C.TH_PASS_DELTA), int(_C.MIN_JFA_PASSES), _pr(_flat["grid"]),
             _pr(_flat["temp"]), _fl.ptr(_info), st)
    majority_filter_labels()
    accumulate_face_area(n)
    finalize_faces_kernel(n, pos, rad, VOXEL_SIZE)
    update_fsc_ema(n, _C.JFA_EMA_ALPHA)
    info = _info.cpu().numpy()
    passes = int(info[0])
    V = JFA_RES ** 3
    delta_history = [float(c) / max(1, V) for c in info[2:2 + passes]]
    fsc_array = fsc.to_numpy()[:n].astype(np.int32)
    fsc_ema_array = fsc_ema.to_numpy()[:n]
    oi = (_fl.I64 * 5)()
    of = (_fl.D * 4)()
    _fl.call("fos_jfa_stats", ctx.handle, n, _pr(rad), voxel, _pr(fsc), _pr(fsc_ema),
             _pr(neighbor_used_count), _pr(face_ids), _pr(face_counts), oi, of, st)
    p_dir = int(oi[0])
    overflow_count = int(oi[1])
    stats = {
        "num_passes": passes,
        "early_exit": bool(info[1]),
        "delta_history": delta_history,
        "jfa_res": JFA_RES,
        "voxel_size": VOXEL_SIZE,
        "mean_score": float(of[2]),
        "std_score": float(of[3]),
        "valid_pairs_pct": 100.0,
        "p_dir": p_dir,
        "p_und_accepted": p_dir // 2,
        "overflow_count": overflow_count,
        "overflow_pct": 100.0 * overflow_count / max(1, n),
        "mean_fsc": float(of[0]),
        "min_fsc": int(oi[2]),
        "max_fsc": int(oi[3]),
        "mean_fsc_ema": float(of[1]),
        "fsc_ema": fsc_ema_array,
    }
    return fsc_array, stats


# ---- telemetry (jfa.py:1204-1330) ------------------------------------------------

def _stats(active_n, rad=None):
    n = int(active_n)
    oi = (_fl.I64 * 5)()
    of = (_fl.D * 4)()
    r = rad if rad is not None else torch.ones(max(1, n), device=_dev())
    _fl.call("fos_jfa_stats", _ctx(n).handle, n, _pr(r), _voxel(), _pr(fsc), _pr(fsc_ema),
             _pr(neighbor_used_count), _pr(face_ids), _pr(face_counts), oi, of, _fl.stream())
    return oi, of


def compute_pair_symmetry(active_n):
    return int(_stats(active_n)[0][0])


def compute_overflow_count(active_n):
    return int(_stats(active_n)[0][1])


def compute_fsc_stats(active_n):
    oi, of = _stats(active_n)
    return (float(of[0]), float(oi[2]), float(oi[3]), float(of[1]))


def compute_score_stats(active_n, rad):
    oi, of = _stats(active_n, rad)
    return (float(of[2]), float(of[3]))


# ---- validation (jfa.py:1337-1426) --------------------------------------------------

def _validate(active_n):
    out = (_fl.I64 * 3)()
    n = int(active_n)
    _fl.call("fos_validate", _ctx(n).handle, n, _pr(neighbor_list), _pr(neighbor_count), out,
             _fl.stream())
    return out


def check_symmetry(active_n):
    return int(_validate(active_n)[0])


def check_overflow(active_n):
    return int(_validate(active_n)[1])


def check_self_loops(active_n):
    return int(_validate(active_n)[2])


def validate_jfa(active_n):
    v = _validate(active_n)
    asym, overflow, self_loops = int(v[0]), int(v[1]), int(v[2])
    return {"passed": overflow == 0 and self_loops == 0, "asymmetric_pairs": asym,
            "overflow_count": overflow, "self_loops": self_loops}


def print_jfa_config():
    """jfa.py:1433-1447."""
    voxel_mem_mb = (JFA_RES ** 3) * 3 * 4 / (1024 ** 2)
    adjacency_mem_mb = MAX_PARTICLES * MAX_NEIGHBORS * 4 / (1024 ** 2)
    print("[JFA] Configuration:")
    print(f"      Resolution: {JFA_RES}³ voxels")
    print(f"      Voxel size: {VOXEL_SIZE:.6f}")
    print(f"      Max neighbors: {MAX_NEIGHBORS}")
    print(f"      Min face voxels: {_C.MIN_FACE_VOXELS}")
    print(f"      Power β: {_C.POWER_BETA}")
    print(f"      Memory: ~{voxel_mem_mb + adjacency_mem_mb * 4:.1f} MB")
