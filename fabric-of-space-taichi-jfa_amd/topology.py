"""Drop-in for the reference topology.py entry points (topology.py:49-504).

The candidate-pair builder is the hash-grid candidate generator of the foam
step; the Gabriel / kNN degree kernels are off in the reference
(TOPO_ENABLED=False, config.py:453) and are provided for completeness.
"""
import torch

import foam_config as _C
import fos_lib as _fl
from fos_lib import Field

topo_deg = None
topo_deg_ema = None
topo_pairs = None
topo_pair_count = None
topo_truncated = None


def allocate_fields(max_n, device="cuda"):
    """topology.py:49-65."""
    global topo_deg, topo_deg_ema, topo_pairs, topo_pair_count, topo_truncated
    d = torch.device(device)
    topo_deg = Field(torch.zeros(max_n, dtype=torch.int32, device=d))
    topo_deg_ema = Field(torch.zeros(max_n, dtype=torch.float32, device=d))
    topo_pairs = Field(torch.zeros(max_n, _C.MAX_TOPO_NEIGHBORS, dtype=torch.int32, device=d))
    topo_pair_count = Field(torch.zeros(max_n, dtype=torch.int32, device=d))
    topo_truncated = Field(torch.zeros(max_n, dtype=torch.int32, device=d))


def build_topo_candidates(pos, rad, cell_start, cell_count, cell_indices, n):
    """topology.py:68-133: j > i within TOPO_MAX_RADIUS_MULTIPLE*(ri+rj)."""
    if topo_pairs is None:
        raise RuntimeError("topology.allocate_fields() must be called first")
    _fl.call("fos_topo_candidates", _fl.ptr(_fl.tensor_of(pos)), _fl.ptr(_fl.tensor_of(rad)),
             int(n), int(_C.GRID_RES), float(_C.DOMAIN_SIZE), float(_C.TOPO_MAX_RADIUS_MULTIPLE),
             int(_C.MAX_TOPO_NEIGHBORS), _fl.ptr(_fl.tensor_of(cell_start)),
             _fl.ptr(_fl.tensor_of(cell_count)), _fl.ptr(_fl.tensor_of(cell_indices)),
             _fl.ptr(topo_pairs), _fl.ptr(topo_pair_count), _fl.ptr(topo_truncated),
             _fl.stream())


def _cs():
    """The configured CELL_SIZE (config.py:290-291); the reference sizes the
    Gabriel stencil with it."""
    return float(_C.CELL_SIZE)


def gabriel_test_topological_degree(pos, rad, cell_start, cell_count, cell_indices, n, frame):
    """topology.py:140-262 over the pairs of build_topo_candidates."""
    _fl.call("fos_topo_gabriel", _fl.ptr(_fl.tensor_of(pos)), int(n), int(_C.GRID_RES),
             float(_C.DOMAIN_SIZE), _cs(), _fl.ptr(_fl.tensor_of(cell_start)),
             _fl.ptr(_fl.tensor_of(cell_count)), _fl.ptr(_fl.tensor_of(cell_indices)),
             _fl.ptr(topo_pairs), _fl.ptr(topo_pair_count), int(_C.MAX_TOPO_NEIGHBORS),
             int(frame), int(_C.TOPO_PAIR_SUBSAMPLING_STRIDE),
             int(bool(_C.TOPO_EARLY_EXIT_HIGH_DEGREE)), float(_C.TOPO_DEG_HIGH),
             _fl.ptr(topo_deg_ema), _fl.ptr(topo_deg), _fl.stream())


def update_topo_ema(ema_alpha, n):
    """topology.py:266-283."""
    _fl.call("fos_topo_ema", _fl.ptr(topo_deg), _fl.ptr(topo_deg_ema), int(n), float(ema_alpha),
             _fl.stream())


def update_radii_topological(rad, deg_effective, deg_low, deg_high, gain_grow, gain_shrink,
                             r_min, r_max, n):
    """topology.py:287-322."""
    _fl.call("fos_topo_radii", _fl.ptr(_fl.tensor_of(rad)), _fl.ptr(_fl.tensor_of(deg_effective)),
             int(n), float(deg_low), float(deg_high), float(gain_grow), float(gain_shrink),
             float(r_min), float(r_max), _fl.stream())


def should_sample_pair(i, j, frame, Q):
    """topology.py:326-333 (host form): 1 for about 1/Q of the pairs."""
    def w(x):
        return ((x + 2 ** 31) % 2 ** 32) - 2 ** 31
    key = w(w(i * 73856093) ^ w(j * 19349663) ^ w(frame * 83492791))
    return 1 if key % Q == 0 else 0


def gabriel_update_batched(pos, rad, cell_start, cell_count, cell_indices, n, batch_idx,
                           num_batches, subsample_q, frame):
    """topology.py:335-446."""
    _fl.call("fos_topo_gabriel_batched", _fl.ptr(_fl.tensor_of(pos)), _fl.ptr(_fl.tensor_of(rad)),
             int(n), int(_C.GRID_RES), float(_C.DOMAIN_SIZE), _cs(),
             float(_C.TOPO_MAX_RADIUS_MULTIPLE), int(_C.MAX_TOPO_NEIGHBORS),
             _fl.ptr(_fl.tensor_of(cell_start)), _fl.ptr(_fl.tensor_of(cell_count)),
             _fl.ptr(_fl.tensor_of(cell_indices)), int(batch_idx), int(num_batches),
             int(subsample_q), int(frame), _fl.ptr(topo_pairs), _fl.ptr(topo_pair_count),
             _fl.ptr(topo_truncated), _fl.ptr(topo_deg), _fl.stream())


def commit_topo_to_ema(alpha, n):
    """topology.py:449-455."""
    update_topo_ema(alpha, n)


def knn_topo_degree(pos, rad, cell_start, cell_count, cell_indices, n, k):
    """topology.py:458-504."""
    _fl.call("fos_topo_knn", _fl.ptr(_fl.tensor_of(pos)), _fl.ptr(_fl.tensor_of(rad)), int(n),
             int(_C.GRID_RES), float(_C.DOMAIN_SIZE), float(_C.TOPO_MAX_RADIUS_MULTIPLE), int(k),
             _fl.ptr(_fl.tensor_of(cell_start)), _fl.ptr(_fl.tensor_of(cell_count)),
             _fl.ptr(_fl.tensor_of(cell_indices)), _fl.ptr(topo_deg), _fl.stream())
