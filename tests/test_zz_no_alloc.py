"""The allocation-free JFA entry point against the context one.

fos_jfa_labels_noalloc keeps its scratch in a static device array; its
labels and pass records must equal fos_jfa_init + fos_jfa_passes bit for
bit, on two calls in a row (the second with a smaller input, so stale
scratch from the first call would show), and it must reject sizes beyond
its static limits.
"""
import ctypes

import numpy as np
import pytest
import torch

import foam_config as fc
import foam_inputs as fi

pytestmark = pytest.mark.gpu


def _with_context(fl, pos, rad, n, res, passes, early):
    ctx = fl.Context(n, res, 27)
    tp = torch.from_numpy(pos).cuda()
    tr = torch.from_numpy(rad).cuda()
    V = res ** 3
    grid = torch.empty(V, dtype=torch.int32, device="cuda")
    temp = torch.empty(V, dtype=torch.int32, device="cuda")
    info = torch.zeros(66, dtype=torch.int32, device="cuda")
    voxel = fc.f32(1.0 / res)
    st = fl.stream()
    fl.call("fos_jfa_init", ctx.handle, fl.ptr(tp), fl.ptr(tr), n, res, 1.0, voxel, 1.0,
            fl.ptr(grid), fl.ptr(temp), st)
    fl.call("fos_jfa_passes", ctx.handle, res, 1.0, voxel, passes, int(early), 0.006, 3,
            fl.ptr(grid), fl.ptr(temp), fl.ptr(info), st)
    torch.cuda.synchronize()
    return grid.cpu().numpy(), info.cpu().numpy()


def _no_alloc(fl, pos, rad, n, res, passes, early):
    tp = torch.from_numpy(pos).cuda()
    tr = torch.from_numpy(rad).cuda()
    V = res ** 3
    # stale values in the outputs must not matter either
    grid = torch.full((V,), 7, dtype=torch.int32, device="cuda")
    temp = torch.full((V,), 5, dtype=torch.int32, device="cuda")
    info = torch.full((66,), 3, dtype=torch.int32, device="cuda")
    voxel = fc.f32(1.0 / res)
    fl.call("fos_jfa_labels_noalloc", fl.ptr(tp), fl.ptr(tr), n, res, 1.0, voxel, 1.0, passes,
            int(early), 0.006, 3, fl.ptr(grid), fl.ptr(temp), fl.ptr(info), fl.stream())
    torch.cuda.synchronize()
    return grid.cpu().numpy(), info.cpu().numpy()


@pytest.mark.parametrize("early", [False, True])
def test_no_alloc_matches_context_entry_point(gpu, early):
    fl = gpu
    cases = [(70000, 128, 0.012, 0.024, 1),   # 32-bit candidate lists (ids > 65,535)
             (3000, 72, 0.02, 0.04, 2)]       # smaller second call, res not a power of 2
    for n, res, lo, hi, seed in cases:
        pos, rad = fi.uniform(n, lo, hi, seed=seed)
        passes = int(np.ceil(np.log2(res))) + 1
        print(f"no-alloc case n={n} res={res} early={early}", flush=True)
        want_g, want_i = _with_context(fl, pos, rad, n, res, passes, early)
        got_g, got_i = _no_alloc(fl, pos, rad, n, res, passes, early)
        assert np.array_equal(got_g, want_g), (n, res)
        k = 2 + passes
        assert np.array_equal(got_i[:k], want_i[:k]), (got_i[:k], want_i[:k])


def test_no_alloc_limits(gpu):
    fl = gpu
    mn, mr, nb = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
    fl.call("fos_static_limits", ctypes.byref(mn), ctypes.byref(mr), ctypes.byref(nb))
    assert mn.value == 262144 and mr.value == 384 and nb.value < 32 * 2 ** 20
    dummy = torch.zeros(16, dtype=torch.float32, device="cuda")
    lab = torch.zeros(16, dtype=torch.int32, device="cuda")
    for n, res in ((mn.value + 1, 8), (8, mr.value + 1)):
        with pytest.raises(fl.FosError, match="static scratch"):
            fl.call("fos_jfa_labels_noalloc", fl.ptr(dummy), fl.ptr(dummy), n, res, 1.0,
                    fc.f32(1.0 / res), 1.0, 4, 0, 0.006, 3, fl.ptr(lab), fl.ptr(lab),
                    fl.ptr(lab), fl.stream())
