"""The oracle reproduces the committed golden fixtures (tests/golden/), and
the benchmark input generators still produce the recorded SHA-256 digests."""
import json
import os

import numpy as np
import pytest

import foam_config as fc
import foam_inputs as fi

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
F = np.float32


@pytest.fixture(scope="module")
def jg():
    return dict(np.load(os.path.join(HERE, "jfa_small.npz")))


@pytest.fixture(scope="module")
def dg():
    return dict(np.load(os.path.join(HERE, "dynamics_small.npz")))


def test_input_digests():
    want = json.load(open(os.path.join(HERE, "inputs_sha256.json")))
    for name, d in want.items():
        pos, rad = fi.for_config(fc.preset(name))
        assert len(rad) == d["n"]
        assert fi.digest(pos) == d["pos"] and fi.digest(rad) == d["rad"], name


def test_oracle_jfa_chain_matches_golden(orc, jg):
    n, res = len(jg["rad"]), round(len(jg["labels"]) ** (1 / 3))
    voxel = fc.f32(1.0 / res)
    seeded, _ = orc.jfa_init(jg["pos"], jg["rad"], n, res, 1.0, voxel)
    assert np.array_equal(seeded, jg["seeded"])
    lab, passes, _, hist = orc.jfa_run(jg["pos"], jg["rad"], n, res, 1.0, voxel, len(jg["hist"]))
    assert np.array_equal(lab, jg["labels"]) and np.array_equal(hist, jg["hist"])
    filt = orc.majority(lab, res)
    assert np.array_equal(filt, jg["filtered"])
    ids, cts = orc.accumulate_faces(filt, res, n)
    assert np.array_equal(ids, jg["face_ids"]) and np.array_equal(cts, jg["face_counts"])
    prev1, nl, nc, used, fsc = orc.finalize_faces(n, jg["rad"], voxel, ids, cts,
                                                  np.zeros((n, 32), F))
    for a, k in ((prev1, "prev1"), (nl, "nl"), (nc, "nc"), (used, "used"), (fsc, "fsc")):
        assert np.array_equal(a, jg[k]), k
    prev2, nl2, nc2, used2, fsc2 = orc.finalize_faces(n, jg["rad2"], voxel, ids, cts, prev1)
    for a, k in ((prev2, "prev2"), (nl2, "nl2"), (nc2, "nc2"), (used2, "used2"), (fsc2, "fsc2")):
        assert np.array_equal(a, jg[k]), k


def test_oracle_dynamics_matches_golden(orc, jg, dg):
    n, G = len(dg["rad"]), round(len(dg["cell_count"]) ** (1 / 3))
    pw, cc, cs, ci = orc.grid_build(dg["pos"], n, G, 1.0)
    for a, k in ((pw, "pos_wrapped"), (cc, "cell_count"), (cs, "cell_start"),
                 (ci, "cell_indices")):
        assert np.array_equal(a, dg[k]), k
    pbd = orc.pbd_jacobi(pw, dg["rad"], n, G, 1.0, fc.f32(1.015), 0.2, 2.0, cs, cc, ci)
    assert np.array_equal(pbd, dg["pbd"])
    maxd, lmd = orc.max_overlap(pw, dg["rad"], n, G, 1.0, cs, cc, ci)
    assert F(maxd) == dg["maxd"] and np.array_equal(lmd, dg["lmd"])
    r = jg["rad"]
    for _ in range(3):
        r, dr, _, _ = orc.pressure(n, 7, r, jg["face_ids"], jg["fsc"], r_min=0.02, r_max=0.08)
    assert np.array_equal(r, dg["pressure_rad"]) and np.array_equal(dr, dg["pressure_dr"])
    assert np.array_equal(orc.brownian(pw, n, F(2e-4), 1.0, 5, 9, 0), dg["brownian_u"])
    assert np.array_equal(orc.brownian(pw, n, F(1e-4), 1.0, 5, 9, 1), dg["brownian_g"])
