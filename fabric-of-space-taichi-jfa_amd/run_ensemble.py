"""Batched ensemble from the command line: B independent foams of the C4
workload (N log-uniform in [500, 20000], JFA edge round(5 N^(1/3)) in
[48, 160]) advanced together, one launch per phase.

    python run_ensemble.py --foams 256 --frames 100 --autotune --graph

--autotune times the JFA pass tile shapes on the batch first; --graph
captures the frame as a hip graph after two eager frames and replays it.
Prints one JSON summary line (batch frames/s, foam-frames/s, mean FSC per
foam: min / median / max).
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import numpy as np  # noqa: E402
import torch  # noqa: E402

import ensemble as en  # noqa: E402


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--foams", type=int, default=256)
    ap.add_argument("--frames", type=int, default=100)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--graph", action="store_true")
    args = ap.parse_args(argv)
    cfgs = en.c4_configs(args.foams, seed=args.seed)
    ens = en.Ensemble(cfgs, [en.c4_inputs(c, b) for b, c in enumerate(cfgs)])
    out = {"foams": args.foams, "particles": int(ens.n), "voxels": int(ens.v0[-1]),
           "frames": args.frames}
    if args.autotune:
        out["jfa_tiles"] = {"picked": None, "ms_per_pass_sequence": ens.autotune_tiles()}
        out["jfa_tiles"]["picked"] = ens.tile_choice
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done = 0
    if args.graph:
        warm = min(2, args.frames)
        for _ in range(warm):
            ens.step()
        ens.capture()
        for _ in range(args.frames - warm):
            ens.replay(check=False)
        ens.check()
    else:
        for _ in range(args.frames):
            ens.step()
    torch.cuda.synchronize()
    done = args.frames
    dt = time.perf_counter() - t0
    fsc = ens.fsc.float()
    means = np.array([float(ens.foam(b, fsc).mean().item()) for b in range(ens.B)])
    out.update({"batch_frames_per_s": round(done / dt, 3),
                "foam_frames_per_s": round(ens.B * done / dt, 1),
                "mean_fsc_per_foam": [round(float(means.min()), 2),
                                      round(float(np.median(means)), 2),
                                      round(float(means.max()), 2)]})
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
