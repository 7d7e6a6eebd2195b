"""numpy/ctypes front end of liboracle.so -- the CPU restatement.

TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the timed CPU
baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg may import this module.  Every function states the serial semantics of
the reference kernel it names (file:line in fos_oracle.c / fos_oracle.h).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
MAXN = 32

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_D = ctypes.c_double
_I64 = ctypes.c_int64
_U64 = ctypes.c_uint64
_U32 = ctypes.c_uint32

_SIG = {
    "orc_set_threads": (None, [_I]),
    "orc_get_threads": (_I, []),
    "orc_power_distance": (_F, [_F, _F, _F, _P, _F, _F, _F]),
    "orc_seed_voxel": (_I, [_P, _I, _F]),
    "orc_jfa_init": (None, [_P, _P, _I, _I, _F, _F, _F, _P, _P]),
    "orc_jfa_step": (_I64, [_P, _P, _I, _F, _F, _F, _I, _P, _P]),
    "orc_jfa_run": (_I, [_P, _P, _I, _I, _F, _F, _F, _I, _I, _D, _I, _P, _P, _P, _P]),
    "orc_majority": (None, [_P, _P, _I]),
    "orc_accumulate_faces": (None, [_P, _I, _I, _P, _P]),
    "orc_finalize_faces": (None, [_I, _P, _F, _P, _P, _P, _P, _P, _P, _P]),
    "orc_fsc_ema": (None, [_I, _F, _P, _P]),
    "orc_jfa_stats": (None, [_I, _P, _F, _P, _P, _P, _P, _P, _P, _P]),
    "orc_validate": (None, [_I, _P, _P, _P]),
    "orc_mark_dirty_tiles": (None, [_P, _P, _P, _P, _I, _I, _F, _F, _I, _I, _F, _F, _P]),
    "orc_grid_build": (None, [_P, _I, _I, _F, _P, _P, _P]),
    "orc_pbd_jacobi": (None, [_P, _P, _P, _I, _I, _F, _F, _F, _F, _P, _P, _P]),
    "orc_pbd_gauss_seidel": (None, [_P, _P, _I, _I, _F, _F, _F, _F, _P, _P, _P]),
    "orc_max_overlap": (_F, [_P, _P, _I, _I, _F, _P, _P, _P, _P]),
    "orc_fsc_targets": (None, [_P, _P, _P, _P, _I, _I, _I, _F, _F, _F, _F]),
    "orc_controller_stats": (None, [_P, _P, _P, _I, _I, _I, _P, _P]),
    "orc_nudge_radii": (None, [_P, _P, _P, _I, _F, _F, _F, _F, _F, _I, _F]),
    "orc_hash": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int32]),
    "orc_pressure": (_F, [_I, _I, _P, _P, _P, _P, _F, _F, _F, _I, _F, _F, _P]),
    "orc_philox": (None, [_U32, _U32, _U32, _U32, _U32, _U32, _P]),
    "orc_brownian": (None, [_P, _I, _F, _F, _U64, _I, _I]),
    "orc_ou_jitter": (None, [_P, _P, _P, _I, _F, _F, _F, _F, _F, _F, _U64, _I]),
    "orc_xsph": (None, [_P, _P, _P, _I, _I, _F, _F, _F, _P, _P, _P]),
    "orc_topo_candidates": (None, [_P, _P, _I, _I, _F, _F, _F, _I, _P, _P, _P, _P, _P, _P]),
    "orc_topo_gabriel": (None, [_P, _I, _I, _F, _F, _P, _P, _P, _P, _P, _I, _I, _I, _I, _F, _P,
                                _P]),
    "orc_topo_gabriel_batched": (None, [_P, _P, _I, _I, _F, _F, _F, _I, _P, _P, _P, _I, _I, _I,
                                        _I, _P, _P, _P, _P]),
    "orc_topo_knn": (None, [_P, _P, _I, _I, _F, _F, _I, _P, _P, _P, _P]),
    "orc_topo_ema": (None, [_P, _P, _I, _F]),
    "orc_topo_radii": (None, [_P, _P, _I, _F, _F, _F, _F, _F, _F]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIG.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_P) if a is not None else None


def f32(a, shape=None):
    out = np.ascontiguousarray(a, dtype=np.float32)
    return out.reshape(shape) if shape is not None else out


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def set_threads(n):
    lib().orc_set_threads(int(n))


def threads():
    return lib().orc_get_threads()


# ---- JFA -----------------------------------------------------------------

def power_distance(c, p, r, beta=1.0, L=1.0):
    p = f32(p, (3,))
    return lib().orc_power_distance(float(c[0]), float(c[1]), float(c[2]), _p(p), r, beta, L)


def seed_voxel(p, res, L=1.0):
    return lib().orc_seed_voxel(_p(f32(p, (3,))), res, L)


def jfa_init(pos, rad, n, res, L, voxel, beta=1.0):
    pos, rad = f32(pos, (-1, 3)), f32(rad)
    V = res ** 3
    grid = np.empty(V, np.int32)
    temp = np.empty(V, np.int32)
    lib().orc_jfa_init(_p(pos), _p(rad), n, res, L, voxel, beta, _p(grid), _p(temp))
    return grid, temp


def jfa_step(pos, rad, res, L, voxel, step, labels_in, beta=1.0):
    pos, rad, lin = f32(pos, (-1, 3)), f32(rad), i32(labels_in)
    out = np.empty_like(lin)
    ch = lib().orc_jfa_step(_p(pos), _p(rad), res, L, voxel, beta, step, _p(lin), _p(out))
    return out, int(ch)


def jfa_run(pos, rad, n, res, L, voxel, num_passes, early_exit=False, th=0.006,
            min_passes=3, beta=1.0):
    """jfa_init + pass loop; returns (labels, passes, early, changes)."""
    pos, rad = f32(pos, (-1, 3)), f32(rad)
    grid, temp = jfa_init(pos, rad, n, res, L, voxel, beta)
    hist = np.zeros(max(1, num_passes), np.int64)
    ee = _I(0)
    p = lib().orc_jfa_run(_p(pos), _p(rad), n, res, L, voxel, beta, num_passes,
                          int(early_exit), th, min_passes, _p(grid), _p(temp), _p(hist),
                          ctypes.byref(ee))
    return grid, int(p), bool(ee.value), hist[:p].copy()


def majority(grid, res):
    g = i32(grid)
    out = np.empty_like(g)
    lib().orc_majority(_p(g), _p(out), res)
    return out


def accumulate_faces(filtered, res, n):
    f = i32(filtered)
    ids = np.empty((n, MAXN), np.int32)
    cts = np.empty((n, MAXN), np.int32)
    lib().orc_accumulate_faces(_p(f), res, n, _p(ids), _p(cts))
    return ids, cts


def finalize_faces(n, rad, voxel, face_ids, face_counts, prev_scores):
    """Returns (prev_scores', neighbor_list, neighbor_count, used, fsc)."""
    rad = f32(rad)
    prev = np.array(prev_scores, np.float32, copy=True).reshape(n, MAXN)
    nl = np.empty((n, MAXN), np.int32)
    nc = np.empty(n, np.int32)
    used = np.empty(n, np.int32)
    fsc = np.empty(n, np.int32)
    lib().orc_finalize_faces(n, _p(rad), voxel, _p(i32(face_ids)), _p(i32(face_counts)),
                             _p(prev), _p(nl), _p(nc), _p(used), _p(fsc))
    return prev, nl, nc, used, fsc


def fsc_ema(n, alpha, fsc, ema):
    e = np.array(ema, np.float32, copy=True)
    lib().orc_fsc_ema(n, alpha, _p(i32(fsc)), _p(e))
    return e


def jfa_stats(n, rad, voxel, fsc, ema, used, face_ids, face_counts):
    oi = np.zeros(5, np.int64)
    of = np.zeros(4, np.float64)
    lib().orc_jfa_stats(n, _p(f32(rad)), voxel, _p(i32(fsc)), _p(f32(ema)), _p(i32(used)),
                        _p(i32(face_ids)), _p(i32(face_counts)), _p(oi), _p(of))
    return oi, of


def validate(n, nl, nc):
    out = np.zeros(3, np.int64)
    lib().orc_validate(n, _p(i32(nl)), _p(i32(nc)), _p(out))
    return out


def mark_dirty_tiles(pos, rad, pos_prev, rad_prev, n, res, L, voxel, tile=16, halo=1,
                     pos_thr=0.5, rad_thr=0.25, dirty=None):
    tpa = (res + tile - 1) // tile
    d = np.zeros(tpa ** 3, np.uint8) if dirty is None else np.array(dirty, np.uint8, copy=True)
    lib().orc_mark_dirty_tiles(_p(f32(pos)), _p(f32(rad)), _p(f32(pos_prev)), _p(f32(rad_prev)),
                               n, res, L, voxel, tile, halo, pos_thr, rad_thr, _p(d))
    return d


# ---- grid / dynamics -------------------------------------------------------

def grid_build(pos, n, G, L):
    """Returns (wrapped pos, cell_count, cell_start, cell_indices)."""
    p = np.array(pos, np.float32, copy=True).reshape(-1, 3)
    C = G ** 3
    cc = np.empty(C, np.int32)
    cs = np.empty(C, np.int32)
    ci = np.empty(max(n, 1), np.int32)
    lib().orc_grid_build(_p(p), n, G, L, _p(cc), _p(cs), _p(ci))
    return p, cc, cs, ci[:n]


def pbd_jacobi(pos, rad, n, G, L, gap_factor, mdf, dmul, cs, cc, ci):
    pin = f32(pos, (-1, 3))
    out = np.empty_like(pin)
    lib().orc_pbd_jacobi(_p(pin), _p(out), _p(f32(rad)), n, G, L, gap_factor, mdf, dmul,
                         _p(i32(cs)), _p(i32(cc)), _p(i32(ci)))
    return out


def pbd_gauss_seidel(pos, rad, n, G, L, gap_factor, mdf, dmul, cs, cc, ci):
    p = np.array(pos, np.float32, copy=True).reshape(-1, 3)
    lib().orc_pbd_gauss_seidel(_p(p), _p(f32(rad)), n, G, L, gap_factor, mdf, dmul,
                               _p(i32(cs)), _p(i32(cc)), _p(i32(ci)))
    return p


def max_overlap(pos, rad, n, G, L, cs, cc, ci):
    lmd = np.zeros(n, np.float32)
    g = lib().orc_max_overlap(_p(f32(pos)), _p(f32(rad)), n, G, L, _p(i32(cs)), _p(i32(cc)),
                              _p(i32(ci)), _p(lmd))
    return float(g), lmd


def fsc_targets(rad, fsc, ema, n, f_low, f_high, growth_pct, r_min, r_max, deadband=1.0):
    rt = np.zeros(n, np.float32)
    lib().orc_fsc_targets(_p(f32(rad)), _p(rt), _p(i32(fsc)), _p(f32(ema)), n, f_low, f_high,
                          growth_pct, r_min, r_max, deadband)
    return rt


def controller_stats(rad, rt, ema, n, f_low, f_high):
    oi = np.zeros(3, np.int64)
    of = np.zeros(2, np.float64)
    lib().orc_controller_stats(_p(f32(rad)), _p(f32(rt)), _p(f32(ema)), n, f_low, f_high,
                               _p(oi), _p(of))
    return oi, of


def nudge_radii(rad, rt, lmd, n, alpha, max_step_pct, r_min, r_max, back_global=1.0,
                mode=2, contact_tol=0.035):
    r = np.array(rad, np.float32, copy=True)
    lib().orc_nudge_radii(_p(r), _p(f32(rt)), _p(f32(lmd)), n, alpha, max_step_pct, r_min,
                          r_max, back_global, mode, contact_tol)
    return r


def lcg_hash(i, frame):
    return int(lib().orc_hash(i, frame))


def pressure(n, frame, rad, face_ids, fsc, k=0.10, p_exp=3.0, cap=0.02, max_nei=10,
             r_min=0.002, r_max=0.010):
    r = np.array(rad, np.float32, copy=True)
    dr = np.zeros(n, np.float32)
    ch = ctypes.c_int32(0)
    mx = lib().orc_pressure(n, frame, _p(r), _p(dr), _p(i32(face_ids)), _p(i32(fsc)), k,
                            p_exp, cap, max_nei, r_min, r_max, ctypes.byref(ch))
    return r, dr, float(mx), int(ch.value)


def philox(c, key):
    out = np.zeros(4, np.uint32)
    lib().orc_philox(c[0], c[1], c[2], c[3], key[0], key[1], _p(out))
    return out


def brownian(pos, n, strength, L, seed, frame, mode):
    p = np.array(pos, np.float32, copy=True).reshape(-1, 3)
    lib().orc_brownian(_p(p), n, strength, L, seed, frame, mode)
    return p


def ou_jitter(pos, v_jit, rad, n, mean_r, dt, rms, tau, drift_cap_frac, L, seed, frame):
    p = np.array(pos, np.float32, copy=True).reshape(-1, 3)
    v = np.array(v_jit, np.float32, copy=True).reshape(-1, 3)
    lib().orc_ou_jitter(_p(p), _p(v), _p(f32(rad)), n, mean_r, dt, rms, tau, drift_cap_frac,
                        L, seed, frame)
    return p, v


def xsph(pos, vel, n, G, L, lim_sq, eps, cs, cc, ci):
    v = np.array(vel, np.float32, copy=True).reshape(-1, 3)
    vt = np.empty_like(v)
    lib().orc_xsph(_p(f32(pos)), _p(v), _p(vt), n, G, L, lim_sq, eps, _p(i32(cs)),
                   _p(i32(cc)), _p(i32(ci)))
    return v


def topo_candidates(pos, rad, n, G, L, mult, cap, cs, cc, ci, r_max=0.0):
    pairs = np.full((n, cap), -1, np.int32)
    cnt = np.zeros(n, np.int32)
    tr = np.zeros(n, np.int32)
    lib().orc_topo_candidates(_p(f32(pos)), _p(f32(rad)), n, G, L, mult, r_max, cap,
                              _p(i32(cs)), _p(i32(cc)), _p(i32(ci)), _p(pairs), _p(cnt), _p(tr))
    return pairs, cnt, tr


# ---- topology.py:140-504 ------------------------------------------------------

def topo_gabriel(pos, n, G, L, cell_size, cs, cc, ci, pairs, pair_count, cap, frame, stride=3,
                 early_exit=True, deg_high=18.0, deg_ema=None):
    deg = np.zeros(n, np.int32)
    ema = np.zeros(n, np.float32) if deg_ema is None else f32(deg_ema)
    lib().orc_topo_gabriel(_p(f32(pos)), n, G, L, cell_size, _p(i32(cs)), _p(i32(cc)),
                           _p(i32(ci)), _p(i32(pairs)), _p(i32(pair_count)), cap, frame, stride,
                           int(early_exit), deg_high, _p(ema), _p(deg))
    return deg


def topo_gabriel_batched(pos, rad, n, G, L, cell_size, mult, cap, cs, cc, ci, batch_idx,
                         num_batches, q, frame, deg=None):
    pairs = np.full((n, cap), -1, np.int32)
    cnt = np.zeros(n, np.int32)
    tr = np.zeros(n, np.int32)
    d = np.zeros(n, np.int32) if deg is None else np.array(deg, np.int32, copy=True)
    lib().orc_topo_gabriel_batched(_p(f32(pos)), _p(f32(rad)), n, G, L, cell_size, mult, cap,
                                   _p(i32(cs)), _p(i32(cc)), _p(i32(ci)), batch_idx, num_batches,
                                   q, frame, _p(pairs), _p(cnt), _p(tr), _p(d))
    return pairs, cnt, tr, d


def topo_knn(pos, rad, n, G, L, mult, k, cs, cc, ci):
    deg = np.zeros(n, np.int32)
    lib().orc_topo_knn(_p(f32(pos)), _p(f32(rad)), n, G, L, mult, k, _p(i32(cs)), _p(i32(cc)),
                       _p(i32(ci)), _p(deg))
    return deg


def topo_ema(deg, ema, n, alpha):
    e = np.array(ema, np.float32, copy=True)
    lib().orc_topo_ema(_p(i32(deg)), _p(e), n, alpha)
    return e


def topo_radii(rad, deg_eff, n, deg_low, deg_high, gain_grow, gain_shrink, r_min, r_max):
    r = np.array(rad, np.float32, copy=True)
    lib().orc_topo_radii(_p(r), _p(f32(deg_eff)), n, deg_low, deg_high, gain_grow, gain_shrink,
                         r_min, r_max)
    return r
