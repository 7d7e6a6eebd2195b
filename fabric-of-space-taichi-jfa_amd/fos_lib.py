"""ctypes binding of libfos.so (the C ABI declared in include/fos.h).

The library is the only compute path: if it is missing or a GPU is absent
the calls raise; there is no CPU fallback.  Tensors cross the boundary as raw
device pointers (``tensor.data_ptr()``) plus sizes, on torch's current HIP
stream.
"""
import ctypes
import os

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfos.so")

P = ctypes.c_void_p
I = ctypes.c_int
U = ctypes.c_uint
I64 = ctypes.c_int64
U64 = ctypes.c_uint64
F = ctypes.c_float
D = ctypes.c_double

# name -> argtypes (every function returns int status)
SIGNATURES = {
    "fos_set_param": [ctypes.c_char_p, I],
    "fos_get_param": [ctypes.c_char_p, ctypes.POINTER(I)],
    "fos_ctx_create": [ctypes.POINTER(P), I, I, I],
    "fos_ctx_destroy": [P],
    "fos_fill_i32": [P, I64, ctypes.c_int32, P],
    "fos_fill_f32": [P, I64, F, P],
    "fos_jfa_seed_table": [P, P, P, I, F, P],
    "fos_jfa_init": [P, P, P, I, I, F, F, F, P, P, P],
    "fos_jfa_step": [P, I, F, F, I, P, P, P, P],
    "fos_jfa_copy": [I, P, P, P],
    "fos_jfa_passes": [P, I, F, F, I, I, D, I, P, P, P, P],
    "fos_jfa_passes_timed": [P, I, F, F, I, I, D, I, P, P, P, P, P],
    "fos_jfa_labels_noalloc": [P, P, I, I, F, F, F, I, I, D, I, P, P, P, P],
    "fos_static_limits": [P, P, P],
    "fos_events_create": [I, P],
    "fos_events_destroy": [I, P],
    "fos_events_elapsed": [I, P, P],
    "fos_majority": [I, P, P, P],
    "fos_face_accumulate": [P, I, I, P, P, P, P],
    "fos_finalize_faces": [P, I, P, F, P, P, P, P, P, P, P, P],
    "fos_fsc_ema": [I, F, P, P, P],
    "fos_jfa_stats": [P, I, P, F, P, P, P, P, P, P, P, P],
    "fos_validate": [P, I, P, P, P, P],
    "fos_dirty_mark": [P, P, P, P, I, I, F, F, I, I, F, F, P, P],
    "fos_dirty_count": [P, P, I, P, P],
    "fos_tile_cache_update": [P, P, I, P, P, P],
    "fos_grid_count": [P, I, I, F, P, P],
    "fos_grid_scan": [P, P, P, I, P],
    "fos_grid_scatter": [P, I, I, F, P, P, P],
    "fos_grid_build": [P, P, I, I, F, P, P, P, P],
    "fos_pbd_pass": [P, P, P, I, I, F, F, F, F, P, P, P, P],
    "fos_max_overlap": [P, P, I, I, F, P, P, P, P, P, P],
    "fos_fsc_targets": [P, P, P, P, I, I, I, F, F, F, F, P],
    "fos_controller_stats": [P, P, P, P, I, I, I, P, P, P],
    "fos_nudge_radii": [P, P, P, I, F, F, F, F, F, I, F, P],
    "fos_pressure_iter": [I, I, P, P, P, P, F, F, F, I, F, F, P, P],
    "fos_pressure_run": [P, I, I, I, P, P, P, P, F, F, F, I, F, F, P, P],
    "fos_pressure_stats": [P, P, I, P, P],
    "fos_pow_third_check": [U, U, U, P, P],
    "fos_brownian": [P, I, F, F, U64, I, I, P],
    "fos_brownian_cells": [P, P, I, F, F, U64, I, I, I, P],
    "fos_settle_nudge": [P, P, P, I, F, F, F, F, I, F, D, P, P, P, I, I, I, P],
    "fos_ou_jitter": [P, P, P, I, F, F, F, F, F, F, U64, I, P],
    "fos_ou_jitter_mean": [P, P, P, P, I, F, F, F, F, F, U64, I, P, P],
    "fos_integrate_jitter": [P, P, I, F, F, P],
    "fos_xsph": [P, P, P, I, I, F, F, F, P, P, P, P],
    "fos_damp": [P, I, F, P],
    "fos_topo_candidates": [P, P, I, I, F, F, I, P, P, P, P, P, P, P],
    "fos_topo_gabriel": [P, I, I, F, F, P, P, P, P, P, I, I, I, I, F, P, P, P],
    "fos_topo_gabriel_batched": [P, P, I, I, F, F, F, I, P, P, P, I, I, I, I, P, P, P, P, P],
    "fos_topo_knn": [P, P, I, I, F, F, I, P, P, P, P, P],
    "fos_topo_ema": [P, P, I, F, P],
    "fos_topo_radii": [P, P, I, F, F, F, F, F, F, P],
    "fos_pbd_pass_cells": [P, P, P, P, I, I, F, F, F, F, P, P, P, P, P],
    "fos_grid_build_from_cells": [P, P, P, I, I, P, P, P, P, P],
    "fos_grid_build_pr": [P, P, P, I, I, F, P, P, P, P, P],
    "fos_jfa_faces": [P, I, I, P, P, F, P, P, P, P, P, P, P, P],
    "fos_jfa_faces_controller": [P, I, I, P, P, P, F, P, P, P, P, P, P, P, P, F, I, I, D, I, I,
                                 F, F, F, F, P, P],
    "fos_fsc_targets_gated": [P, P, P, P, I, I, I, F, F, F, F, P, P],
    "fos_fsc_gate": [P, P, I, I, I, D, P, P],
    "fos_batch_create": [P, I, P, P, P, P, P, P, P, P, P, F],
    "fos_batch_destroy": [P],
    "fos_batch_info": [P, P],
    "fos_batch_set_tiles": [P, P],
    "fos_batch_jfa_init": [P, P, P, P, F, F, P, P, P],
    "fos_batch_jfa_passes": [P, P, F, I, D, I, P, P, P, P],
    "fos_batch_majority": [P, P, P, P],
    "fos_batch_face_accumulate": [P, P, P, P, P, I, P],
    "fos_batch_finalize_faces": [P, P, P, P, P, P, P, P, P, P, I, P],
    "fos_batch_flags": [P, P, P],
    "fos_batch_control_pre": [P, P, P, I, I, D, D, P, P, P, P, P],
    "fos_batch_control_post": [P, P, P, I, D, I, I, P, P, P],
    "fos_batch_frame_advance": [P, P],
    "fos_batch_grid_build": [P, P, P, F, P, P, P, P, I, P, P, P],
    "fos_batch_max_overlap": [P, P, P, F, P, P, P, P, P, P, P],
    "fos_batch_pbd_pass": [P, P, P, P, F, F, F, F, P, P, P, P, P, I, P],
    "fos_batch_nudge_radii": [P, P, P, P, F, F, I, F, P, P, P],
    "fos_batch_fsc_targets": [P, P, P, P, P, F, F, P, P],
    "fos_batch_pressure_run": [P, P, I, P, I, P, P, P, P, F, F, F, I, P],
    "fos_batch_brownian": [P, P, F, F, I, P, I, P],
    "fos_batch_count_zero": [P, P, P, P],
    "fos_pred_candidates": [P, P, I, I, F, F, I, P, P, P, P, P, P, P, P, P],
    "fos_pred_pairs": [P, P, I, P, P, P],
    "fos_pred_features": [P, I, F, F, P, P, P, P, P, I, I64, P, P, P],
    "fos_pred_prev_rows": [P, P, I, P, P],
    "fos_pred_mlp": [P, I64, P, P, P, P, P, F, P, P, P],
    "fos_pred_fsc": [P, P, P, I, I, P, P, P, P],
    "fos_pred_labels": [P, P, P, I, I, I64, P, P, P, P],
    "fos_mlp_param_count": [ctypes.POINTER(I)],
    "fos_mlp_train_step": [P, P, P, P, I, F, P, P, P, I, F, P, P, P],
    "fos_frame_step": [P, P, P, P, P, P],
    "fos_frame_run": [P, P, P, P, I, P, P],
    "fos_frame_release": [P],
    "fos_sched_stats": [P, P, P, P, I, F, I, I, P, P, I, P, P],
    "fos_jfa_schedule": [P, P, P, P, P, P, P, P],
}


class FrameDesc(ctypes.Structure):
    """fos_frame_desc (include/fos.h): the frame's configuration."""
    _fields_ = [("n", I), ("res", I), ("G", I), ("jfa_passes", I),
                ("L", F), ("voxel", F), ("beta", F), ("ema_alpha", F),
                ("early_exit", I), ("min_passes", I), ("th_pass_delta", D),
                ("pbd_base", I), ("pbd_max", I), ("pbd_scale", D), ("r_max_d", D),
                ("gap_factor", F), ("max_disp_frac", F), ("disp_mult", F),
                ("alpha", F), ("max_step_pct", F), ("r_min", F), ("r_max", F),
                ("contact_tol", F), ("contact_tol_d", D),
                ("backpressure_mode", I), ("adjustment_frames", I), ("warmstart_frames", I),
                ("fsc_zero_rate", D), ("f_low", I), ("f_high", I),
                ("growth_pct", F), ("deadband", F),
                ("pressure", I), ("pressure_iters", I), ("pressure_max_nei", I),
                ("pressure_k", F), ("pressure_exp", F), ("pressure_cap", F),
                ("brownian", I), ("brownian_mode", I), ("brownian_strength", F),
                ("seed", U64), ("ou_jitter", I),
                ("dt", F), ("jitter_rms", F), ("jitter_tau", F), ("drift_cap_frac", F),
                ("jfa_every", I), ("jfa_warm", I), ("slot_records", I),
                ("schedule", I), ("skip_warmstart", I), ("max_skip", I),
                ("th_rsigma", D), ("th_margin", D), ("th_fsc_drift", D), ("oob_frac", D),
                ("adaptive_res", I), ("res_min", I), ("res_max", I),
                ("voxels_per_diameter", D), ("L_d", D)]


class FrameBufs(ctypes.Structure):
    """fos_frame_bufs (include/fos.h): device pointers of the frame's arrays."""
    _fields_ = [("pos", P * 2)] + [(k, P) for k in (
        "rad", "rad_target", "lmd", "delta_r", "v_jit", "cell_count", "cell_start",
        "cell_indices", "cell_pr", "grid", "temp", "filtered", "face_ids", "face_counts", "fsc",
        "fsc_ema", "neighbor_list", "neighbor_count", "used_count", "prev_scores", "info", "maxd",
        "maxd_host", "settle_dev", "pos_prev", "rad_prev", "sched_host")]


class FrameState(ctypes.Structure):
    """fos_frame_state (include/fos.h): what carries from frame to frame."""
    _fields_ = [("frame", I), ("pos_cur", I), ("sidx", I), ("settle_force", I),
                ("settle_gate", I), ("need_rebuild", I), ("pending", I),
                ("pbd_passes", I), ("jfa_ran", I), ("gate_read", I),
                ("max_depth", F), ("gate", F), ("event", P),
                ("sched_frames", I), ("since", I), ("have_prev", I), ("fsc_mu_set", I),
                ("fsc_mu_pending", I), ("lockout", I), ("reason", I), ("jfa_res", I),
                ("last_fsc_mu", D), ("signals", D * 5)]

# reasons of a JFA scheduling decision (FOS_JFA_* in include/fos.h)
JFA_REASONS = {0: "skip", 1: "warm-start", 2: "band-change", 3: "watchdog",
               4: "out-of-band-guard", 5: "activity", 6: "cadence"}

_lib = None


class FosError(RuntimeError):
    pass


def load():
    """Loads libfos.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FosError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    for name, argtypes in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = ctypes.c_int
    lib.fos_abi_version.restype = ctypes.c_int
    lib.fos_abi_version.argtypes = []
    lib.fos_last_error.restype = ctypes.c_char_p
    lib.fos_last_error.argtypes = []
    _lib = lib
    return lib


def call(name, *args):
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.fos_last_error().decode(errors="replace")
        raise FosError(f"{name} failed ({rc}): {msg}")
    return rc


def require_gpu():
    if not torch.cuda.is_available():
        raise FosError("libfos needs a ROCm GPU; torch.cuda.is_available() is False")


def stream():
    """Handle of torch's current HIP stream."""
    return torch.cuda.current_stream().cuda_stream


def ptr(t):
    """Device pointer of a contiguous tensor (None -> NULL)."""
    if t is None:
        return None
    if isinstance(t, Field):
        t = t.tensor
    if not t.is_cuda:
        raise FosError("expected a device tensor")
    if not t.is_contiguous():
        raise FosError("expected a contiguous tensor")
    return t.data_ptr()


def tensor_of(x):
    return x.tensor if isinstance(x, Field) else x


class Context:
    """Scratch owned by libfos for up to max_n particles, max_res voxels per
    axis and max_cells hash cells."""

    def __init__(self, max_n, max_res, max_cells):
        require_gpu()
        handle = P()
        call("fos_ctx_create", ctypes.byref(handle), int(max_n), int(max_res), int(max_cells))
        self.handle = handle
        self.max_n, self.max_res, self.max_cells = int(max_n), int(max_res), int(max_cells)

    def fits(self, n, res, cells):
        return n <= self.max_n and res <= self.max_res and cells <= self.max_cells

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib is not None:
            try:
                _lib.fos_ctx_destroy(h)
            except Exception:
                pass
            self.handle = None


_ctx = None


def context(n=1, res=1, cells=27):
    """Process-wide context, grown when a call needs more room."""
    global _ctx
    if _ctx is None or not _ctx.fits(n, res, cells):
        old = _ctx
        mn = max(n, old.max_n if old else 1)
        mr = max(res, old.max_res if old else 1)
        mc = max(cells, old.max_cells if old else 27)
        if old is not None:
            torch.cuda.synchronize()
        _ctx = Context(mn, mr, mc)
    return _ctx


class Field:
    """Taichi-field look-alike over a torch device tensor: ``to_numpy()``,
    ``from_numpy()``, ``fill()``, indexing (``[None]`` for 0-D fields)."""

    def __init__(self, tensor):
        self.tensor = tensor

    @property
    def shape(self):
        return tuple(self.tensor.shape)

    @property
    def dtype(self):
        return self.tensor.dtype

    def to_numpy(self):
        return self.tensor.detach().cpu().numpy()

    def to_torch(self):
        return self.tensor

    def from_numpy(self, arr):
        self.tensor.copy_(torch.from_numpy(np.ascontiguousarray(arr)).to(self.tensor.device))

    def fill(self, value):
        self.tensor.fill_(value)

    def _key(self, key):
        return () if key is None else key

    def __getitem__(self, key):
        v = self.tensor[self._key(key)]
        return v.item() if v.dim() == 0 else v

    def __setitem__(self, key, value):
        self.tensor[self._key(key)] = value

    def __len__(self):
        return self.tensor.shape[0]

    def __repr__(self):
        return f"Field(shape={self.shape}, dtype={self.dtype})"
