"""Telemetry line formats of the reference frame loop (run.py), so logs of
this build read like the reference's and the same parsers work on them.

Each function returns the lines run.py prints, as strings, from the stats
dicts of jfa.run_jfa / jfa.validate_jfa (same keys) or plain numbers.
"""
import re

import numpy as np


def jfa_step4_lines(stats, validation, t_jfa, res, n):
    """run.py:1095-1149: the per-JFA summary (three lines and a status line).

    stats: the dict jfa.run_jfa returns; validation: jfa.validate_jfa's dict;
    t_jfa: seconds spent in the JFA; res: JFA resolution; n: active particles.
    """
    mean_fsc = stats.get("mean_fsc", 0.0)
    min_fsc = stats.get("min_fsc", 0)
    max_fsc = stats.get("max_fsc", 0)
    mean_fsc_ema = stats.get("mean_fsc_ema", 0.0)
    mean_score = stats.get("mean_score", 0.0)
    std_score = stats.get("std_score", 0.0)
    overflow_pct = stats.get("overflow_pct", 0.0)
    p_dir = stats.get("p_dir", 0)
    p_und = stats.get("p_und_accepted", 0)
    asym_pct = 100.0 * validation["asymmetric_pairs"] / max(1, n)
    lines = [f"[JFA Step 4] FSC μ={mean_fsc:.1f} [{min_fsc},{max_fsc}] | "
             f"EMA μ={mean_fsc_ema:.1f}",
             f"      score: μ={mean_score:.2f} σ={std_score:.2f} | "
             f"pairs: {p_und} (P_dir={p_dir})"]
    hist = stats.get("delta_history", [])
    if hist:
        if stats["early_exit"]:
            parts = [f"{d * 100:.1f}%" for d in hist[:-1]] + [f"{hist[-1] * 100:.1f}%*"]
        else:
            parts = [f"{d * 100:.1f}%" for d in hist]
        delta = f" | δ=[{', '.join(parts[:6])}]"
    else:
        delta = ""
    max_passes = int(np.ceil(np.log2(res))) + 1
    lines.append(f"      asym={asym_pct:.1f}% overflow={overflow_pct:.1f}% | "
                 f"time={t_jfa * 1000:.1f}ms | res={res}³ | "
                 f"passes={stats['num_passes']}/{max_passes} "
                 f"(early_exit={stats['early_exit']}){delta}")
    ok = (asym_pct < 1.0 and overflow_pct == 0.0 and 0.8 <= mean_score <= 1.2
          and mean_fsc >= 10.0)
    if validation["passed"] and ok:
        lines.append("      ✅ Step 4 SUCCESS - All metrics in target range")
    elif validation["passed"]:
        lines.append("      ✓ Validation passed (still converging)")
    else:
        lines.append("      ⚠️  Validation FAILED")
    return lines


def jfa_schedule_lines(rec, sched_frames, warmstart_frames, th_rsigma, th_margin_frac,
                       th_fsc_drift, max_skip):
    """run.py:988-1015: the temporal-skip decision lines.

    rec: a frame record with jfa_reason, jfa_since and jfa_signals (FoamSim's
    schedule_record()); sched_frames: the scheduler's frame count (1 on the
    first decision); the rest the configuration's thresholds.
    """
    reason, since, sig = rec["jfa_reason"], rec["jfa_since"], rec["jfa_signals"]
    if reason != "skip":
        lines = []
        if reason == "warm-start" and sched_frames == 1:
            lines.append(f"[JFA Skip] Warm-start: running every frame for first "
                         f"{warmstart_frames} frames")
        if reason == "out-of-band-guard":
            reason = f"out-of-band-guard ({sig['out_of_band'] * 100:.1f}%)"
        lines.append(f"[JFA Run ] reason={reason} | since={since}")
        return lines
    th_margin = th_margin_frac * sig["r_mean"]
    return [f"[JFA Skip] σ(Δr)/r={sig['sigma_dr_over_r']:.4f} ≤ {th_rsigma:.4f} | "
            f"vmax*dt={sig['vmax_dt']:.6f} ≤ {th_margin:.6f} | "
            f"ΔμFSC={sig['delta_fsc_mu']:.3f} ≤ {th_fsc_drift:.3f} | "
            f"since={since}/{max_skip}"]


def dirty_tiles_line(dirty, total, halo, pos_thr, rad_thr, res, L=1.0):
    """run.py:1156-1163: the dirty-tile instrumentation line."""
    pct = 100.0 * dirty / max(1, total)
    voxel = L / res
    return (f"[JFA][PhaseA] tiles={dirty}/{total} ({pct:.1f}%) | halo={halo} | "
            f"Δpos_thr={pos_thr}×{voxel:.6f} | Δr_thr={rad_thr}×{voxel:.6f}")


def frame_lines(frame, n, passes, max_depth, fsc, rad, f_low, f_high, rescue_pct=0.0,
                rescue_enabled=False):
    """run.py:1303-1318: the every-100-frames summary (fsc, rad: host arrays
    of the active particles)."""
    fsc = np.asarray(fsc)
    rad = np.asarray(rad)
    fm = float(fsc.mean()) if n > 0 else 0.0
    fmin = int(fsc.min()) if n > 0 else 0
    fmax = int(fsc.max()) if n > 0 else 0
    rescue = f"{rescue_pct:.1f}%" if rescue_enabled else f"{rescue_pct:.1f}% (disabled)"
    return [f"[Frame {frame:4d}] N={n}, Passes={passes}, MaxDepth={max_depth:.6f}",
            f"    FSC: mean={fm:.1f}, min={fmin}, max={fmax} | band=[{f_low},{f_high}]",
            f"    Radius: mean={rad.mean():.4f}, min={rad.min():.4f}, max={rad.max():.4f}",
            f"    Rescue: {rescue}"]


def parse_step4(lines):
    """Reads back the numbers of jfa_step4_lines (for log tooling): mean/min/
    max FSC, EMA mean, score mean/std, pairs, P_dir, passes, max passes."""
    a = re.match(r"\[JFA Step 4\] FSC μ=([-\d.]+) \[(-?\d+),(-?\d+)\] \| EMA μ=([-\d.]+)",
                 lines[0])
    b = re.match(r"\s+score: μ=([-\d.]+) σ=([-\d.]+) \| pairs: (\d+) \(P_dir=(\d+)\)",
                 lines[1])
    c = re.search(r"passes=(\d+)/(\d+) \(early_exit=(True|False)\)", lines[2])
    if not (a and b and c):
        raise ValueError("not a JFA Step 4 block")
    return {"mean_fsc": float(a.group(1)), "min_fsc": int(a.group(2)),
            "max_fsc": int(a.group(3)), "mean_fsc_ema": float(a.group(4)),
            "mean_score": float(b.group(1)), "std_score": float(b.group(2)),
            "p_und_accepted": int(b.group(3)), "p_dir": int(b.group(4)),
            "num_passes": int(c.group(1)), "max_passes": int(c.group(2)),
            "early_exit": c.group(3) == "True"}
