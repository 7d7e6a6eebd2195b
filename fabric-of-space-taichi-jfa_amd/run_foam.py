"""Headless frame driver from the command line (the run.py:700-1261 frame
loop without the GUI).

    python run_foam.py --config c2 --frames 100
    python run_foam.py --config c5s --frames 48 --predictor face_mlp_c5.npz
    python run_foam.py --config c1 --frames 300 --telemetry --csv foam.csv

Prints the reference-format telemetry lines with --telemetry, and one JSON
summary line at the end (frames/s over the run, judge errors when the
learned face predictor is on).
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

import foam_config as fc  # noqa: E402
import foam_inputs as fi  # noqa: E402
from foam import FoamSim  # noqa: E402


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="c2", help="preset name (foam_config.preset)")
    ap.add_argument("--frames", type=int, default=None, help="default: the preset's frames")
    ap.add_argument("--predictor", default=None,
                    help="weights .npz of the learned face predictor (presets with judge_every)")
    ap.add_argument("--judge-on", choices=("per_particle", "mean"), default=None,
                    help="what the predictor's judge bounds (default: the preset's judge_on)")
    ap.add_argument("--telemetry", action="store_true")
    ap.add_argument("--csv", default=None, help="write ID,X,Y,Z,Radius,FSC at the end")
    ap.add_argument("--save", default=None, help="write the state snapshot (.npz) at the end")
    args = ap.parse_args(argv)
    cfg = fc.preset(args.config)
    if args.judge_on:
        cfg.judge_on = args.judge_on
    frames = cfg.frames if args.frames is None else args.frames
    pos, rad = fi.for_config(cfg)
    sim = FoamSim(cfg, pos, rad, telemetry=args.telemetry)
    if args.predictor:
        if cfg.judge_every <= 0:
            raise SystemExit(f"preset {args.config} has no judge cadence (judge_every = 0)")
        w = np.load(args.predictor, allow_pickle=False)
        sim.set_predictor(w["params"], json.loads(str(w["meta"]))["mult"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if args.telemetry:
        for _ in range(frames):
            sim.step()
            for line in sim.log_lines():
                print(line, flush=True)
    else:
        sim.run(frames)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    if args.csv:
        sim.export_csv(args.csv)
    if args.save:
        sim.save(args.save)
    out = {"config": args.config, "n": cfg.n, "frames": frames,
           "frames_per_s": round(frames / dt, 3) if dt > 0 else None,
           "mean_fsc": float(sim.fsc.float().mean().item())}
    if args.predictor:
        out["judged"] = [{"frame": f, "fsc_err_per_particle": round(e, 5),
                          "fsc_err_mean_fsc": round(b, 5)} for f, e, b in sim.judged]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
