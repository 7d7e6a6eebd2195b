"""The HIP path against the committed golden fixtures (tests/golden/),
through the C ABI, with no oracle at run time."""
import os

import numpy as np
import pytest
import torch

import foam_config as fc

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
F = np.float32
_alive = []


@pytest.fixture(autouse=True)
def _keep():
    yield
    torch.cuda.synchronize()
    _alive.clear()


def dev(a):
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    _alive.append(t)
    return t


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def empty(shape, dtype):
    t = torch.empty(shape, dtype=dtype, device="cuda")
    _alive.append(t)
    return t


def test_jfa_chain_matches_golden(gpu):
    g = dict(np.load(os.path.join(HERE, "jfa_small.npz")))
    n, res = len(g["rad"]), round(len(g["labels"]) ** (1 / 3))
    P = len(g["hist"])
    voxel = fc.f32(1.0 / res)
    print(f"[golden jfa] n={n} res={res} passes={P}", flush=True)
    ctx = gpu.Context(n, res, 27)
    st, p = gpu.stream(), gpu.ptr
    tp, tr = dev(g["pos"]), dev(g["rad"])
    grid, temp = empty(res ** 3, torch.int32), empty(res ** 3, torch.int32)
    info = torch.zeros(2 + P, dtype=torch.int32, device="cuda")
    gpu.call("fos_jfa_init", ctx.handle, p(tp), p(tr), n, res, 1.0, voxel, 1.0, p(grid), p(temp),
             st)
    assert np.array_equal(host(grid), g["seeded"])
    gpu.call("fos_jfa_passes", ctx.handle, res, 1.0, voxel, P, 0, 0.006, 3, p(grid), p(temp),
             p(info), st)
    assert np.array_equal(host(grid), g["labels"])
    assert np.array_equal(host(info)[2:], g["hist"])
    filt = empty(res ** 3, torch.int32)
    gpu.call("fos_majority", res, p(grid), p(filt), st)
    assert np.array_equal(host(filt), g["filtered"])
    ids, cts = empty((n, 32), torch.int32), empty((n, 32), torch.int32)
    gpu.call("fos_face_accumulate", ctx.handle, res, n, p(filt), p(ids), p(cts), st)
    assert np.array_equal(host(ids), g["face_ids"]) and np.array_equal(host(cts), g["face_counts"])
    prev = torch.zeros((n, 32), dtype=torch.float32, device="cuda")
    nl, nc = empty((n, 32), torch.int32), empty(n, torch.int32)
    used, fsc = empty(n, torch.int32), empty(n, torch.int32)
    for suffix, rad in (("", g["rad"]), ("2", g["rad2"])):
        gpu.call("fos_finalize_faces", ctx.handle, n, p(dev(rad)), voxel, p(ids), p(cts),
                 p(prev), p(nl), p(nc), p(used), p(fsc), st)
        k = "prev1" if suffix == "" else "prev2"
        assert np.array_equal(host(prev), g[k]), k
        for t, name in ((nl, "nl"), (nc, "nc"), (used, "used"), (fsc, "fsc")):
            assert np.array_equal(host(t), g[name + suffix]), name + suffix


def test_dynamics_match_golden(gpu):
    d = dict(np.load(os.path.join(HERE, "dynamics_small.npz")))
    j = dict(np.load(os.path.join(HERE, "jfa_small.npz")))
    n, G = len(d["rad"]), round(len(d["cell_count"]) ** (1 / 3))
    print(f"[golden dynamics] n={n} G={G}", flush=True)
    ctx = gpu.Context(n, 1, G ** 3)
    st, p = gpu.stream(), gpu.ptr
    tp, tr = dev(d["pos"]), dev(d["rad"])
    cc, cs = empty(G ** 3, torch.int32), empty(G ** 3, torch.int32)
    ci = empty(n, torch.int32)
    gpu.call("fos_grid_build", ctx.handle, p(tp), n, G, 1.0, p(cc), p(cs), p(ci), st)
    for t, k in ((tp, "pos_wrapped"), (cc, "cell_count"), (cs, "cell_start"),
                 (ci, "cell_indices")):
        assert np.array_equal(host(t), d[k]), k
    out = torch.empty_like(tp)
    gpu.call("fos_pbd_pass", p(tp), p(out), p(tr), n, G, 1.0, fc.f32(1.015), 0.2, 2.0, p(cs),
             p(cc), p(ci), st)
    assert np.array_equal(host(out), d["pbd"])
    lmd, mx = empty(n, torch.float32), empty(1, torch.float32)
    gpu.call("fos_max_overlap", p(tp), p(tr), n, G, 1.0, p(cs), p(cc), p(ci), p(lmd), p(mx), st)
    assert np.array_equal(host(lmd), d["lmd"]) and F(host(mx)[0]) == d["maxd"]
    rad = dev(j["rad"])
    dr = empty(n, torch.float32)
    gpu.call("fos_pressure_run", ctx.handle, n, 7, 3, p(rad), p(dr), p(dev(j["face_ids"])),
             p(dev(j["fsc"])), 0.1, 3.0, 0.02, 10, 0.02, 0.08, None, st)
    assert np.array_equal(host(rad), d["pressure_rad"])
    assert np.array_equal(host(dr), d["pressure_dr"])
    tb = dev(d["pos_wrapped"])
    gpu.call("fos_brownian", p(tb), n, 2e-4, 1.0, 5, 9, 0, st)
    assert np.array_equal(host(tb), d["brownian_u"])
    tb = dev(d["pos_wrapped"])
    gpu.call("fos_brownian", p(tb), n, 1e-4, 1.0, 5, 9, 1, st)
    # Box-Muller goes through double log/cos, whose last bit may differ
    # between the device and host libraries before the rounding to fp32
    assert np.allclose(host(tb), d["brownian_g"], rtol=0, atol=1e-9)
