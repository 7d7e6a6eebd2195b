"""Property tests of the HIP kernels against the oracle over random small
inputs (hypothesis; one process, bounded example counts):

* JFA seeding and passes on random particle sets, including coincident
  particles (seed and distance ties) and non-power-of-two grids;
* majority filter, face accumulation and finalize on random label grids
  with random hysteresis state (integer results bit-exact);
* grid rebuild and one PBD pass on random positions, partly outside the box.
"""
import math

import numpy as np
import pytest
import torch
from hypothesis import given, settings
from hypothesis import strategies as st

import foam_config as fc
from test_gpu_parity import G, dev, gpu_grid, host
from test_oracle_props import random_labels

pytestmark = pytest.mark.gpu
F = np.float32
SETTINGS = settings(max_examples=20, deadline=None)


@SETTINGS
@given(seed=st.integers(0, 2 ** 31 - 1), n=st.integers(1, 300), res=st.integers(5, 24),
       dup=st.booleans())
def test_jfa_random_particles(gpu, orc, seed, n, res, dup):
    rng = np.random.default_rng(seed)
    pos = (rng.random((n, 3)) - 0.5).astype(F)
    rad = rng.uniform(0.005, 0.08, n).astype(F)
    if dup and n > 1:
        pos[1::2] = pos[0::2][: n // 2]
        rad[1::2] = rad[0::2][: n // 2]
    voxel = fc.f32(1.0 / res)
    P = int(math.ceil(math.log2(res))) + 1
    lab, passes, _, hist = orc.jfa_run(pos, rad, n, res, 1.0, voxel, P)
    g = G(gpu, n, res)
    grid = torch.empty(res ** 3, dtype=torch.int32, device="cuda")
    temp = torch.empty_like(grid)
    info = torch.zeros(66, dtype=torch.int32, device="cuda")
    st_ = gpu.stream()
    g.call("fos_jfa_init", g.h, g.ptr(dev(pos)), g.ptr(dev(rad)), n, res, 1.0, voxel, 1.0,
           g.ptr(grid), g.ptr(temp), st_)
    g.call("fos_jfa_passes", g.h, res, 1.0, voxel, P, 0, 0.006, 3, g.ptr(grid), g.ptr(temp),
           g.ptr(info), st_)
    assert np.array_equal(host(grid), lab)
    inf = host(info)
    assert inf[0] == passes and (inf[2:2 + passes] == hist).all()


@SETTINGS
@given(seed=st.integers(0, 2 ** 31 - 1), res=st.integers(3, 12), n=st.integers(1, 40),
       holes=st.sampled_from([0.0, 0.1]))
def test_faces_and_finalize_random_labels(gpu, orc, seed, res, n, holes):
    rng = np.random.default_rng(seed)
    lab = random_labels(seed, res, n, holes)
    rad = rng.uniform(0.01, 0.06, n).astype(F)
    prev = ((rng.random((n, 32)) < 0.4) * rng.uniform(0.05, 0.3, (n, 32))).astype(F)
    voxel = fc.f32(1.0 / res)
    filt = orc.majority(lab, res)
    ids, cts = orc.accumulate_faces(filt, res, n)
    want = orc.finalize_faces(n, rad, voxel, ids, cts, prev)
    g = G(gpu, n, res)
    st_ = gpu.stream()
    tf = torch.empty(res ** 3, dtype=torch.int32, device="cuda")
    g.call("fos_majority", res, g.ptr(dev(lab)), g.ptr(tf), st_)
    assert np.array_equal(host(tf), filt)
    fid = torch.empty(n, 32, dtype=torch.int32, device="cuda")
    fct = torch.empty(n, 32, dtype=torch.int32, device="cuda")
    tp = dev(prev)
    out = [torch.empty(n, 32, dtype=torch.int32, device="cuda")] + \
        [torch.empty(n, dtype=torch.int32, device="cuda") for _ in range(3)]
    g.call("fos_jfa_faces", g.h, res, n, g.ptr(tf), g.ptr(dev(rad)), voxel, g.ptr(fid),
           g.ptr(fct), g.ptr(tp), *[g.ptr(t) for t in out], st_)
    assert np.array_equal(host(fid), ids) and np.array_equal(host(fct), cts)
    got = [host(tp)] + [host(t) for t in out]
    for a, b, name in zip(want, got, ("prev", "nl", "nc", "used", "fsc")):
        assert np.array_equal(a, b), name


@SETTINGS
@given(seed=st.integers(0, 2 ** 31 - 1), n=st.integers(1, 600), G_=st.integers(3, 12),
       spread=st.sampled_from([0.5, 0.8]))
def test_grid_and_pbd_random_positions(gpu, orc, seed, n, G_, spread):
    rng = np.random.default_rng(seed)
    pos = ((rng.random((n, 3)) - 0.5) * 2 * spread).astype(F)
    rad = rng.uniform(0.01, 0.05, n).astype(F)
    p, cc, cs, ci = orc.grid_build(pos, n, G_, 1.0)
    tp, tcc, tcs, tci = gpu_grid(gpu, pos, n, G_)
    assert np.array_equal(host(tp), p) and np.array_equal(host(tci), ci)
    assert np.array_equal(host(tcc), cc) and np.array_equal(host(tcs), cs)
    gap = fc.f32(1.015)
    want = orc.pbd_jacobi(p, rad, n, G_, 1.0, gap, 0.2, 2.0, cs, cc, ci)
    out = torch.empty_like(tp)
    gpu.call("fos_pbd_pass", gpu.ptr(tp), gpu.ptr(out), gpu.ptr(dev(rad)), n, G_, 1.0, gap, 0.2,
             2.0, gpu.ptr(tcs), gpu.ptr(tcc), gpu.ptr(tci), gpu.stream())
    assert np.array_equal(host(out), want)
