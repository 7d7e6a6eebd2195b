"""Drop-in for the reference grid.py: periodic CSR spatial hash.

The five rebuild kernels keep their names and argument lists
(grid.py:145-309); fields are torch device tensors (or ``fos_lib.Field``).
GRID_RES, DOMAIN_SIZE, HALF_L and INV_L are read from foam_config at call
time.  The scan is a real device-wide exclusive scan and the scatter lists
each cell's particles in ascending index -- the serial semantics of the
reference loops, independent of thread timing.
"""
import foam_config as _C
import fos_lib as _fl


def _pr(x):
    return _fl.ptr(_fl.tensor_of(x))


def _G():
    return int(_C.GRID_RES)


def _L():
    return float(_C.DOMAIN_SIZE)


def clear_grid(cell_count):
    """grid.py:145-153."""
    _fl.call("fos_fill_i32", _pr(cell_count), _G() ** 3, 0, _fl.stream())


def count_particles_per_cell(pos, cell_count, n):
    """grid.py:189-222."""
    _fl.call("fos_grid_count", _pr(pos), int(n), _G(), _L(), _pr(cell_count), _fl.stream())


def prefix_sum(cell_count, cell_start):
    """grid.py:229-243."""
    G = _G()
    ctx = _fl.context(1, 1, G ** 3)
    _fl.call("fos_grid_scan", ctx.handle, _pr(cell_count), _pr(cell_start), G, _fl.stream())


def copy_cell_pointers(cell_start, cell_write):
    """grid.py:250-261."""
    _fl.tensor_of(cell_write)[: _G() ** 3].copy_(_fl.tensor_of(cell_start)[: _G() ** 3])


def scatter_particles(pos, cell_write, cell_indices, n):
    """grid.py:268-309 (stores wrapped positions back into pos)."""
    _fl.call("fos_grid_scatter", _pr(pos), int(n), _G(), _L(), _pr(cell_write),
             _pr(cell_indices), _fl.stream())


def rebuild(pos, n, cell_count, cell_start, cell_indices):
    """run.py:476-504 rebuild_grid in one call (clear, count, scan, copy,
    scatter)."""
    G = _G()
    ctx = _fl.context(int(n), 1, G ** 3)
    _fl.call("fos_grid_build", ctx.handle, _pr(pos), int(n), G, _L(), _pr(cell_count),
             _pr(cell_start), _pr(cell_indices), _fl.stream())


def update_colors(deg, color, n, deg_low, deg_high):
    """grid.py:316 is visualisation only (out of scope)."""
    raise NotImplementedError("update_colors is GUI colouring, not part of the foam step")
