"""Drop-in for the reference dynamics.py hot-path kernels.

Names, parameters and return types follow dynamics.py (keyword names too:
run.py calls equilibrate_pressure and apply_brownian_motion with keywords).
Constants come from foam_config at call time.  Deliberate differences:

* project_overlaps is the Jacobi form: every particle reads the pass-start
  positions (the reference's parallel loop reads positions other threads are
  writing; its serial order is index-order Gauss-Seidel).
* reductions are true reductions (reduce_max_depth, the pressure and
  controller statistics), not the reference's racy plain assignments.
* noise comes from Philox4x32-10 keyed by (RNG_SEED, call counter, particle,
  axis) instead of ti.random; ``frame`` arguments select the counter.
* powers in equilibrate_pressure are evaluated in double and rounded once.
"""
import torch

import foam_config as _C
import fos_lib as _fl

_noise_calls = {"brownian": 0, "ou": 0}


def _pr(x):
    return _fl.ptr(_fl.tensor_of(x))


def _G():
    return int(_C.GRID_RES)


def _L():
    return float(_C.DOMAIN_SIZE)


def _t(x):
    return _fl.tensor_of(x)


# ---- dynamics.py:77-171 ------------------------------------------------------

def project_overlaps(pos, rad, cell_start, cell_count, cell_indices, n):
    n = int(n)
    p = _t(pos)
    out = torch.empty_like(p)
    _fl.call("fos_pbd_pass", _pr(p), _pr(out), _pr(rad), n, _G(), _L(),
             _C.f32(1.0 + _C.GAP_FRACTION), _C.MAX_DISPLACEMENT_FRAC,
             _C.DISPLACEMENT_MULTIPLIER, _pr(cell_start), _pr(cell_count), _pr(cell_indices),
             _fl.stream())
    p[:n].copy_(out[:n])


# ---- dynamics.py:304-372 ------------------------------------------------------

_maxd = None


def compute_local_max_overlaps(pos, rad, cell_start, cell_count, cell_indices,
                               local_max_depth, n):
    global _maxd
    if _maxd is None or _maxd.device != _t(pos).device:
        _maxd = torch.zeros(1, dtype=torch.float32, device=_t(pos).device)
    _fl.call("fos_max_overlap", _pr(pos), _pr(rad), int(n), _G(), _L(), _pr(cell_start),
             _pr(cell_count), _pr(cell_indices), _pr(local_max_depth), _pr(_maxd),
             _fl.stream())


def reduce_max_depth(local_max_depth, n):
    t = _t(local_max_depth)[: int(n)]
    return float(t.max().item()) if int(n) > 0 else 0.0


def compute_max_overlap(pos, rad, cell_start, cell_count, cell_indices, local_max_depth, n):
    compute_local_max_overlaps(pos, rad, cell_start, cell_count, cell_indices,
                               local_max_depth, n)
    return float(_maxd.item())


# ---- dynamics.py:473-480, 532-593 (velocity paths) -------------------------------

def apply_global_damping(vel, n):
    _fl.call("fos_damp", _pr(vel), int(n), _C.GLOBAL_DAMPING, _fl.stream())


def apply_xsph_smoothing(pos, vel, vel_temp, rad, cell_start, cell_count, cell_indices, n):
    lim = _C.f32((2.0 * _C.R_MAX) ** 2)
    _fl.call("fos_xsph", _pr(pos), _pr(vel), _pr(vel_temp), int(n), _G(), _L(), lim,
             _C.XSPH_EPSILON, _pr(cell_start), _pr(cell_count), _pr(cell_indices),
             _fl.stream())


# ---- dynamics.py:600-660 OU jitter ---------------------------------------------

def init_jitter_velocities(v_jit, n):
    _t(v_jit)[: int(n)].zero_()


def apply_brownian(v_jit, rad, mean_radius, n, dt):
    """OU velocity update; positions move in integrate_jitter."""
    if not _C.JITTER_ENABLED:
        return
    mr = mean_radius[None] if hasattr(mean_radius, "__getitem__") else float(mean_radius)
    k = _noise_calls["ou"]
    _noise_calls["ou"] = k + 1
    _fl.call("fos_ou_jitter", None, _pr(v_jit), _pr(rad), int(n), float(mr), float(dt),
             _C.JITTER_RMS, _C.JITTER_TAU, _C.f32(_C.MAX_DRIFT_FRACTION * _C.GAP_FRACTION),
             _L(), int(_C.RNG_SEED), k, _fl.stream())


def integrate_jitter(pos, v_jit, n, dt):
    if not _C.JITTER_ENABLED:
        return
    _fl.call("fos_integrate_jitter", _pr(pos), _pr(v_jit), int(n), float(dt), _L(),
             _fl.stream())


# ---- dynamics.py:1226-1427 FSC controller ------------------------------------------

def set_fsc_targets(rad, rad_target, fsc, fsc_ema, n, f_low, f_high, growth_pct, r_min, r_max):
    _fl.call("fos_fsc_targets", _pr(rad), _pr(rad_target), _pr(fsc), _pr(fsc_ema), int(n),
             int(f_low), int(f_high), float(growth_pct), float(r_min), float(r_max),
             float(_C.FSC_DEADBAND), _fl.stream())


def compute_controller_stats(rad, rad_target, fsc_ema, n, f_low, f_high):
    oi = (_fl.I64 * 3)()
    of = (_fl.D * 2)()
    _fl.call("fos_controller_stats", _fl.context().handle, _pr(rad), _pr(rad_target),
             _pr(fsc_ema), int(n), int(f_low), int(f_high), oi, of, _fl.stream())
    return int(oi[0]), int(oi[1]), int(oi[2]), float(of[0]), float(of[1])


def nudge_radii_adaptive_ema(rad, rad_target, local_max_depth, n, alpha, max_step_pct,
                             r_min, r_max, back_global, mode):
    _fl.call("fos_nudge_radii", _pr(rad), _pr(rad_target), _pr(local_max_depth), int(n),
             float(alpha), float(max_step_pct), float(r_min), float(r_max), float(back_global),
             int(mode), float(_C.CONTACT_TOL), _fl.stream())


# ---- dynamics.py:1436-1589 pressure --------------------------------------------------

def hash_func(i, frame):
    """dynamics.py:1436-1451 (int32 wraparound)."""
    t = (int(i) + 12345 * int(frame)) & 0xFFFFFFFF
    return ((1103515245 * t + 12345) & 0xFFFFFFFF) & 0x7FFFFFFF


_pstats = None


def equilibrate_pressure(n, frame, pos, rad, delta_r, jfa_face_ids, jfa_fsc, k, P_exp,
                         pair_cap, max_nei, r_min, r_max):
    """Returns (max_abs_dr, changed_count)."""
    global _pstats
    dev = _t(rad).device
    if _pstats is None or _pstats.device != dev:
        _pstats = torch.zeros(2, dtype=torch.float32, device=dev)
    _fl.call("fos_pressure_iter", int(n), int(frame), _pr(rad), _pr(delta_r), _pr(jfa_face_ids),
             _pr(jfa_fsc), float(k), float(P_exp), float(pair_cap), int(max_nei), float(r_min),
             float(r_max), _pr(_pstats), _fl.stream())
    s = _pstats.cpu()
    return float(s[0].item()), int(s.view(torch.int32)[1].item())


def compute_pressure_stats(n, rad, P_exp, r_min, r_max):
    out = (_fl.F * 2)()
    _fl.call("fos_pressure_stats", _fl.context().handle, _pr(rad), int(n), out, _fl.stream())
    return float(out[0]), float(out[1])


# ---- dynamics.py:1596-1634 ---------------------------------------------------------------

def apply_brownian_motion(pos, n, strength, domain_size, frame=None, gaussian=False):
    """Uniform displacement (u-0.5)*2*strength per axis, wrapped by one L.
    ``frame`` keys the noise (default: a per-process call counter);
    ``gaussian=True`` draws N(0, strength^2) instead."""
    if frame is None:
        frame = _noise_calls["brownian"]
        _noise_calls["brownian"] += 1
    _fl.call("fos_brownian", _pr(pos), int(n), float(strength), float(domain_size),
             int(_C.RNG_SEED), int(frame), 1 if gaussian else 0, _fl.stream())
