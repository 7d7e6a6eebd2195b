"""Writes the golden fixtures under tests/golden/ from the oracle.

    python tests/golden/make_golden.py

The reference ships no golden vectors (its tests hold prose expectations
only), so these fixtures are outputs of the oracle -- the C restatement of
the reference's serial semantics, itself pinned by the closed-form,
brute-force and known-answer tests in tests/test_oracle_*.py -- on small
seeded inputs.  They freeze the oracle (tests/test_golden.py) and give the
GPU path fixed expectations that need no oracle at run time
(tests/test_gpu_golden.py).  inputs_sha256.json records the SHA-256 of the
benchmark inputs (C1, C2, C3 generators with seed 0).
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "fabric-of-space-taichi-jfa_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import foam_config as fc      # noqa: E402
import foam_inputs as fi      # noqa: E402
from oracle import oracle as O  # noqa: E402

F = np.float32
N, RES, G = 300, 24, 7
PASSES = 6                    # ceil(log2 24) + 1


def jfa_case():
    pos, rad = fi.uniform(N, 0.03, 0.07, seed=21)
    voxel = fc.f32(1.0 / RES)
    seeded, _ = O.jfa_init(pos, rad, N, RES, 1.0, voxel)
    labels, passes, _, hist = O.jfa_run(pos, rad, N, RES, 1.0, voxel, PASSES)
    filt = O.majority(labels, RES)
    ids, cts = O.accumulate_faces(filt, RES, N)
    prev0 = np.zeros((N, 32), F)
    prev1, nl, nc, used, fsc = O.finalize_faces(N, rad, voxel, ids, cts, prev0)
    # a second frame with grown radii exercises the hysteresis on prev1
    rad2 = (rad * F(1.03)).astype(F)
    prev2, nl2, nc2, used2, fsc2 = O.finalize_faces(N, rad2, voxel, ids, cts, prev1)
    return dict(pos=pos, rad=rad, rad2=rad2, seeded=seeded, labels=labels,
                hist=np.asarray(hist, np.int64), filtered=filt, face_ids=ids, face_counts=cts,
                prev1=prev1, nl=nl, nc=nc, used=used, fsc=fsc, prev2=prev2, nl2=nl2, nc2=nc2,
                used2=used2, fsc2=fsc2)


def dynamics_case(j):
    pos, rad = fi.uniform(N, 0.03, 0.07, seed=22)
    pos[::4] += F(0.8)                              # unwrapped inputs for the grid
    pw, cc, cs, ci = O.grid_build(pos, N, G, 1.0)
    gap = fc.f32(1.015)
    pbd = O.pbd_jacobi(pw, rad, N, G, 1.0, gap, 0.2, 2.0, cs, cc, ci)
    maxd, lmd = O.max_overlap(pw, rad, N, G, 1.0, cs, cc, ci)
    r = j["rad"]
    for it in range(3):
        r, dr, mx, ch = O.pressure(N, 7, r, j["face_ids"], j["fsc"], r_min=0.02, r_max=0.08)
    bu = O.brownian(pw, N, F(2e-4), 1.0, 5, 9, 0)
    bg = O.brownian(pw, N, F(1e-4), 1.0, 5, 9, 1)
    return dict(pos=pos, rad=rad, pos_wrapped=pw, cell_count=cc, cell_start=cs,
                cell_indices=ci, pbd=pbd, maxd=np.float32(maxd), lmd=lmd,
                pressure_rad=r, pressure_dr=dr, brownian_u=bu, brownian_g=bg)


def input_digests():
    out = {}
    for name in ("c1", "c2", "c3"):
        pos, rad = fi.for_config(fc.preset(name))
        out[name] = {"pos": fi.digest(pos), "rad": fi.digest(rad), "n": int(len(rad))}
    return out


def main():
    j = jfa_case()
    np.savez_compressed(os.path.join(HERE, "jfa_small.npz"), **j)
    d = dynamics_case(j)
    np.savez_compressed(os.path.join(HERE, "dynamics_small.npz"), **d)
    with open(os.path.join(HERE, "inputs_sha256.json"), "w") as f:
        json.dump(input_digests(), f, indent=1, sort_keys=True)
        f.write("\n")
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
