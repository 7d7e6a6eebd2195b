"""Drop-in for the reference topology.py entry points.

Only the candidate-pair builder is part of the foam step's hot path (it is
the hash-grid candidate generator); the Gabriel / kNN degree kernels are off
in the reference (TOPO_ENABLED=False, config.py:453) and are not provided.
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


def _off(name):
    def f(*args, **kwargs):
        raise NotImplementedError(f"topology.{name} is disabled in the reference "
                                  "(TOPO_ENABLED=False) and not provided")
    f.__name__ = name
    return f


gabriel_test_topological_degree = _off("gabriel_test_topological_degree")
update_topo_ema = _off("update_topo_ema")
update_radii_topological = _off("update_radii_topological")
gabriel_update_batched = _off("gabriel_update_batched")
commit_topo_to_ema = _off("commit_topo_to_ema")
knn_topo_degree = _off("knn_topo_degree")
