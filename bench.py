"""Benchmark: per-frame foam step at N=10,000, JFA 128^3 (one MI355X).

A step is one frame of the foam (hash grid, PBD, pressure, noise, JFA power
diagram, FSC extraction, controller) on the uniform-box workload: N=10,000,
radii U[0.012, 0.024], JFA 128^3 int32 labels, 8 passes, JFA every frame.
Inputs are resident in HBM before the timed region.

The JSON line also carries:
  roofline      the JFA pass kernel: algorithmic 8 B per voxel-pass (read and
                write one int32 label) times 128^3 voxels per launch, over the
                average launch duration from HIP events on its stream; traffic
                is the HBM bytes per launch that rocprofv3 counted
                (FETCH_SIZE + WRITE_SIZE, each scaled by what the counter
                reads for a plain copy of known size), read from the profile
                summary under profiles/ when it is there
  cpu_baseline  the CPU restatement (C, OpenMP over the host cores) timed on
                a bounded sample of the same frames
  cadence       the same workload with JFA on a 1/5 cadence after 20 frames
  c3_jfa        the JFA alone on the large clustered foam (N=200,000, 384^3)
  c3_frames     whole frames of that foam
  surrogate     the learned face predictor at N=50,000 / 256^3: frames/s with
                the JFA every frame against the JFA every 8th frame (the
                judge) and the predictor in between, with the judged FSC
                errors (the frame's mean FSC, and per particle)
  ensemble      the batched ensemble: 256 independent foams (N log-uniform in
                [500, 20000], JFA edge round(5 N^(1/3)) in [48, 160]) advanced
                by one launch per phase; foam-frames/s of the batch against
                the same foams stepped one after another (a sample of them)
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "fabric-of-space-taichi-jfa_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "frames/s at N=10k,128³ JFA; Gvox-pass/s + %HBM; surrogate vs plain @FSC≤2%"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_jfa_pass_c2_hbm_traffic.md")
WEIGHTS = os.path.join(PKG, "face_mlp_c5.npz")
COPY_CAL_RES = 400          # 400^3 int32 = 256,000,000 B read and written


JFA_SRC = os.path.join(PKG, "csrc", "fos_jfa.hip")
COMMON_SRC = os.path.join(PKG, "csrc", "fos_common.h")


def kernel_signature():
    """SHA-256 (first 16 hex digits) of the JFA pass kernel's source -- the
    regions of fos_jfa.hip between its pass-kernel and pass-launch markers,
    the shared device helpers and the build flags -- so a counter profile can
    be matched to the kernel the bench times."""
    import hashlib
    import fos_build
    src = open(JFA_SRC).read()
    h = hashlib.sha256()
    for tag in ("kernel", "launch"):
        a = src.index(f"// @jfa-pass-{tag}-begin")
        b = src.index(f"// @jfa-pass-{tag}-end")
        h.update(src[a:b].encode())
    h.update(open(COMMON_SRC).read().encode())
    h.update(" ".join(fos_build.FLAGS).encode())
    return h.hexdigest()[:16]


def measured_traffic(path=TRAFFIC_FILE):
    """(HBM bytes per pass launch, kernel signature) from a committed counter
    summary, or (None, None)."""
    traffic = sig = None
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    traffic = int(float(line.split(":", 1)[1]))
                elif line.startswith("kernel_signature:"):
                    sig = line.split(":", 1)[1].strip()
    except OSError:
        pass
    return traffic, sig


def traffic_fields(path):
    """roofline.traffic for the timed kernel: the profile's figure when the
    profile was taken on this kernel source, else null with the reason."""
    traffic, sig = measured_traffic(path)
    now = kernel_signature()
    if traffic is not None and sig == now:
        return {"traffic": traffic, "kernel_signature": now}
    return {"traffic": None, "kernel_signature": now,
            "traffic_note": (f"{os.path.relpath(path, ROOT)} was taken on kernel signature {sig}"
                             if traffic is not None else "no counter profile")}


PMC_FILE = os.path.join(ROOT, "profiles", "k_jfa_pass_list_c2_pmc.md")
# wave64 VALU instructions the chip can issue per second: 256 CUs x 4 SIMDs,
# one instruction per SIMD every 4 cycles (16 lanes), 2.4 GHz
VALU_PEAK_WAVE_INSTS = 256 * 4 * 2.4e9 / 4


def valu_fields(avg_launch_s, path=PMC_FILE):
    """The pass kernel against its instruction-issue ceiling: the profile's
    mean VALU instructions per launch over the bench's own mean launch time,
    when the profile was taken on this kernel source."""
    insts = sig = None
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("valu_wave_insts_per_launch:"):
                    insts = float(line.split(":", 1)[1])
                elif line.startswith("kernel_signature:"):
                    sig = line.split(":", 1)[1].strip()
    except OSError:
        pass
    if insts is None or sig != kernel_signature():
        return {"valu_issue": None}
    rate = insts / avg_launch_s
    return {"valu_issue": {"achieved": round(rate / 1e9, 1), "peak": round(VALU_PEAK_WAVE_INSTS / 1e9, 1),
                           "unit": "G wave64 VALU inst/s", "frac": round(rate / VALU_PEAK_WAVE_INSTS, 3),
                           "source": os.path.relpath(path, ROOT)}}


def copy_calibration():
    """Three launches of a plain int32 copy of COPY_CAL_RES^3 elements, for
    scaling the FETCH_SIZE / WRITE_SIZE counters (run under rocprofv3 --pmc)."""
    import torch

    import fos_lib as fl
    n = COPY_CAL_RES ** 3
    src = torch.arange(n, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        fl.call("fos_jfa_copy", COPY_CAL_RES, fl.ptr(src), fl.ptr(dst), fl.stream())
    torch.cuda.synchronize()
    assert torch.equal(src, dst)
    del src, dst
    torch.cuda.empty_cache()


def copy_bandwidth(reps=10):
    """Achieved HBM bandwidth of a plain int32 copy of COPY_CAL_RES^3 elements
    (256,000,000 B read and 256,000,000 B written per launch), timed with HIP
    events: the attainable ceiling beside the 8 TB/s peak."""
    import torch

    import fos_lib as fl
    n = COPY_CAL_RES ** 3
    src = torch.arange(n, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    fl.call("fos_jfa_copy", COPY_CAL_RES, fl.ptr(src), fl.ptr(dst), fl.stream())
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fl.call("fos_jfa_copy", COPY_CAL_RES, fl.ptr(src), fl.ptr(dst), fl.stream())
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    torch.cuda.empty_cache()
    return round(2.0 * 4 * n / (ms / 1e3) / 1e9, 1)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-frames", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cadence", action="store_true")
    ap.add_argument("--no-temporal", action="store_true")
    ap.add_argument("--no-early-exit", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--surrogate-steps", type=int, default=48)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-foams", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=20)
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--c3-frames", type=int, default=4)
    ap.add_argument("--pressure-per-launch", action="store_true",
                    help="one pressure kernel per iteration (fos_set_param pressure_flow 0)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain copy of 400^3 int32 (counter scaling)")
    return ap.parse_args()


def run_frames(sim, frames):
    sim.run(frames)


EVENT_EVERY = 8   # timed frames between two with HIP events around every JFA pass


def timed(cfg, pos, rad, steps, warmup, with_events, setup=None):
    """Wall time of `steps` frames after `warmup`.  with_events: on every
    EVENT_EVERY-th timed frame HIP events bracket each JFA pass (the per-pass
    times), and halfway between those frames two events bracket the whole pass
    schedule (the mean launch time without a record between the passes; each
    record costs the stream a few microseconds).  Returns (seconds, per-pass
    ms per instrumented frame, whole-schedule ms per bracketed frame, sim)."""
    import torch

    import fos_lib as fl
    from foam import FoamSim
    sim = FoamSim(cfg, pos, rad)
    if setup is not None:
        setup(sim)
    run_frames(sim, warmup)
    torch.cuda.synchronize()
    P = cfg.passes()
    ev_arrays, ends = [], []
    if with_events:
        slots = []
        for k in range(steps):
            arr = None
            if k % EVENT_EVERY == 0:
                arr = (ctypes.c_void_p * (P + 1))()
                fl.call("fos_events_create", P + 1, arr)
                ev_arrays.append(arr)
            elif k % EVENT_EVERY == EVENT_EVERY // 2:
                two = (ctypes.c_void_p * 2)()
                fl.call("fos_events_create", 2, two)
                arr = (ctypes.c_void_p * (P + 1))()
                arr[0], arr[P] = two[0], two[1]
                ends.append(two)
            slots.append(arr)
        sim.pass_events = slots
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run_frames(sim, steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    pass_ms, sched_ms = [], []
    for arr in ev_arrays:
        ms = (ctypes.c_float * P)()
        fl.call("fos_events_elapsed", P + 1, arr, ms)
        pass_ms.append(list(ms))
        fl.call("fos_events_destroy", P + 1, arr)
    for two in ends:
        ms = (ctypes.c_float * 1)()
        fl.call("fos_events_elapsed", 2, two, ms)
        sched_ms.append(float(ms[0]))
        fl.call("fos_events_destroy", 2, two)
    return t1 - t0, pass_ms, sched_ms, sim


def temporal_bench(cfg, pos, rad, steps, warmup):
    """C2 under the reference's temporal-skip scheduler (run.py:902-1018):
    frames/s past the scheduler's warm start and why each timed frame ran
    or skipped the JFA (one host read per frame for the decision)."""
    import collections

    import torch

    from foam import FoamSim
    sim = FoamSim(cfg, pos, rad)
    sim.run(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reasons = [sim.step()["jfa_reason"] for _ in range(steps)]
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    return {"workload": "same, JFA on the frames the temporal-skip scheduler picks "
                        "(TH_RSIGMA 0.004, TH_MARGIN_SAFE 0.4, TH_FSC_DRIFT 0.05, "
                        "MAX_SKIP_FRAMES 15, 30-frame warm start)",
            "value": round(steps / el, 3), "unit": "frames/s",
            "frames": [warmup, warmup + steps - 1],
            "jfa_frames": sum(1 for r in reasons if r != "skip"),
            "reasons": dict(collections.Counter(reasons))}


def early_exit_bench(cfg, pos, rad, steps, warmup, sample=20):
    """C2 with the reference's default early exit inside the pass schedule
    (config.py:208-210: stop once fewer than 0.6% of the voxels changed in a
    pass, after at least 3 passes; the device skips the remaining passes):
    frames/s over the timed window, then the passes the JFA ran on `sample`
    further frames, read back one frame at a time outside the timed window."""
    import collections
    import dataclasses

    import torch

    from foam import FoamSim
    c = dataclasses.replace(cfg, early_exit=True)
    sim = FoamSim(c, pos, rad)
    sim.run(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.run(steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    passes = []
    for _ in range(sample):
        sim.step()
        passes.append(int(sim.info[0].item()))
    return {"workload": f"same, early exit on (threshold {c.th_pass_delta}, at least "
                        f"{c.min_passes} passes; the reference default, config.py:208-210)",
            "value": round(steps / el, 3), "unit": "frames/s",
            "frames": [warmup, warmup + steps - 1],
            "passes_run": {str(k): v for k, v in sorted(collections.Counter(passes).items())},
            "passes_sample": f"frames {warmup + steps}..{warmup + steps + sample - 1}"}


C3_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_jfa_pass_c3_hbm_traffic.md")


def c3_jfa_bench(frames):
    """The JFA alone at C3 (N=200,000 clustered, 384^3, 10 passes): the label
    ping-pong (2 x 226 MB) exceeds the 256 MB Infinity Cache, so this is the
    size where the passes stream from HBM.  Each frame re-seeds and runs the
    whole pass schedule with HIP events around every pass."""
    import numpy as np
    import torch

    import foam_config as fc
    import foam_inputs as fi
    import fos_lib as fl
    from foam import FoamSim
    cfg = fc.preset("c3")
    pos, rad = fi.for_config(cfg)
    sim = FoamSim(cfg, pos, rad)
    c, st, P, V = cfg, fl.stream(), cfg.passes(), cfg.res ** 3
    voxel = c.voxel()
    ms_all = []
    for k in range(frames + 1):
        arr = (ctypes.c_void_p * (P + 1))()
        fl.call("fos_events_create", P + 1, arr)
        fl.call("fos_jfa_init", sim.ctx.handle, fl.ptr(sim.pos), fl.ptr(sim.rad), sim.n, sim.res,
                c.L, voxel, c.beta, fl.ptr(sim.grid), None, st)
        fl.call("fos_jfa_passes_timed", sim.ctx.handle, sim.res, c.L, voxel, P,
                int(c.early_exit), c.th_pass_delta, c.min_passes, fl.ptr(sim.grid),
                fl.ptr(sim.temp), fl.ptr(sim.info), arr, st)
        ms = (ctypes.c_float * P)()
        fl.call("fos_events_elapsed", P + 1, arr, ms)
        fl.call("fos_events_destroy", P + 1, arr)
        if k > 0:                      # the first run warms up
            ms_all.append(list(ms))
    torch.cuda.synchronize()
    flat = np.array(ms_all, dtype=np.float64)
    avg_pass_s = float(flat.mean()) / 1e3
    achieved = 8.0 * V / avg_pass_s / 1e9
    return {"workload": "C3 JFA: N=200,000 clustered (64 centres, sigma 0.05, plus 100,000 uniform), "
                        "radii lognormal in [0.003, 0.015], 384^3 int32 labels, 10 passes",
            "frames": frames, "jfa_ms": round(float(flat.sum(axis=1).mean()), 3),
            "gvox_pass_per_s": round(P * V / (float(flat.sum(axis=1).mean()) / 1e3) / 1e9, 2),
            "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                         **traffic_fields(C3_TRAFFIC_FILE),
                         "kernel": "k_jfa_pass_list", "avg_launch_us": round(avg_pass_s * 1e6, 2),
                         "per_pass_us": [round(x * 1e3, 1) for x in flat.mean(axis=0)]}}


def c3_frames_bench(frames, warmup=2):
    """Whole C3 frames (N=200,000 clustered, 384^3, the run.py:700-1261
    order): frames/s of FoamSim on the large foam, inputs resident."""
    import foam_config as fc
    import foam_inputs as fi
    cfg = fc.preset("c3")
    pos, rad = fi.for_config(cfg)
    el, _, _, sim = timed(cfg, pos, rad, frames, warmup, False)
    return {"workload": (f"C3 frames: N={cfg.n:,} clustered, radii lognormal clipped to "
                         f"[0.003, 0.015], JFA {cfg.res}^3 every frame, PBD {cfg.pbd_base}.."
                         f"{cfg.pbd_max}, hash {cfg.hash_res()}^3, pressure {cfg.pressure_iters} "
                         f"iteration"),
            "frames": frames, "frames_per_s": round(frames / el, 3),
            "ms_per_frame": round(1e3 * el / frames, 3),
            "pbd_passes_last": int(sim.last["pbd_passes"])}


def candidate_coverage(sim, mult):
    """Share of the last JFA frame's neighbour relations (both directions)
    that the predictor's candidate rows at `mult` (ri + rj) hold, with the
    candidates per particle and the rows cut at the row capacity."""
    import torch

    import face_predictor as fp
    n = sim.cfg.n
    pr = fp.FacePredictor(fp.init_params(0), n, sim.cfg.L, sim.cfg.r_max, mult=mult)
    pr.candidates(sim.pos, sim.rad, n)
    lab = torch.empty(n * pr.cap, dtype=torch.float32, device="cuda")
    pr.labels(n, sim.neighbor_list, sim.neighbor_count, lab)
    P = pr.npairs(n)
    found = float(lab[:P].sum().item())
    rel = float(sim.neighbor_count[:n].sum().item())
    return {"coverage": round(2.0 * found / max(1.0, rel), 5),
            "candidates_per_particle": round(float(pr.ncand[:n].float().mean().item()), 2),
            "truncated_rows": int(pr.trunc[:n].sum().item())}


def surrogate_bench(steps, warmup):
    """Learned face predictor on the N=50,000 / 256^3 workload: frames/s of the
    plain path (JFA every frame) and of the predictor path (JFA every 8th
    frame as the judge, predictor in between, fallback to the JFA while the
    judged FSC error exceeds 2%), and the judged errors."""
    import json as _json

    import numpy as np

    import foam_config as fc
    import foam_inputs as fi
    w = np.load(WEIGHTS, allow_pickle=False)
    params, meta = w["params"], _json.loads(str(w["meta"]))
    plain = fc.preset("c5")
    pos, rad = fi.for_config(plain)
    el_p, _, _, psim = timed(plain, pos, rad, steps, warmup, False)
    fp_ = steps / el_p
    cover = {str(m): candidate_coverage(psim, m) for m in (1.6, meta["mult"])}
    out = {"workload": "N=50000, radii U[0.005,0.011], JFA 256^3; predictor path: JFA every "
                       "8th frame as judge, BF16 MLP on MFMA in between, JFA on every frame "
                       "while the judged error exceeds 2%",
           "plain_fps": round(fp_, 3), "frames": steps,
           "candidate_mult": meta["mult"],
           "candidate_coverage": cover,
           "trained_on": (f"{meta['frames']} plain-JFA frames of {len(meta['train_seeds'])} "
                          f"foams (input seeds {meta['train_seeds']}); the benchmark's foam "
                          f"(seed {meta['holdout_seed']}) only validates"),
           "features": 13,
           "departures": (
               "candidates within 3.5 (ri + rj), not 1.6: the radii fill ~14% of the box, so "
               "faces reach far beyond the spheres; candidate_coverage is the share of the "
               "JFA's neighbour relations the candidate rows hold at each multiple (1.6 caps "
               "the per-particle error far above 2% before any scoring). 13 features, not 12: "
               "the 13th is the previous frame's acceptance, which the JFA's face hysteresis "
               "uses (jfa.py:400-443). 240 training frames, not 20,000: 32 frames of one foam "
               "and 240 frames of six foams give the same validation error (2.91%), so more "
               "frames of this workload do not lower it")}
    # the judge bounds either every particle's FSC (per_particle) or the
    # frame's mean FSC (mean, where over- and under-counts cancel)
    for judge_on in ("per_particle", "mean"):
        surr = fc.preset("c5s")
        surr.judge_on = judge_on
        el_s, _, _, sim = timed(surr, pos, rad, steps, warmup, False,
                             setup=lambda s: s.set_predictor(params, meta["mult"]))
        judged = [j for j in sim.judged if j[0] >= warmup]
        k = 1 if judge_on == "per_particle" else 2
        fs_ = steps / el_s
        out[f"judge_{judge_on}"] = {
            "surrogate_fps": round(fs_, 3), "speedup": round(fs_ / fp_, 3),
            "judge_frames": len(judged),
            "jfa_frames": int(sum(1 for f in range(warmup, warmup + steps)
                                  if f in {j[0] for j in sim.judged})),
            "fsc_err_per_particle": (round(float(np.mean([j[1] for j in judged])), 5)
                                     if judged else None),
            "fsc_err_mean_fsc": (round(float(np.mean([j[2] for j in judged])), 5)
                                 if judged else None),
            "fallback_judgements": int(sum(1 for j in judged if j[k] > surr.max_fsc_err))}
    out["note"] = ("the <=2% bound holds for the frame's mean FSC only (judge_mean); the "
                   "per-particle error is about 3% (1% of it the relations the candidates miss), "
                   "so judged per particle at 2% the path falls back to the JFA on every frame "
                   "(judge_per_particle)")
    return out


def ensemble_bench(B, steps, warmup, sample=8):
    """Batched ensemble of B foams: frames of the whole batch per second and
    foam-frames/s, against a sample of the same foams stepped one at a time
    (FoamSim), whose mean frame time gives the one-after-another rate."""
    import numpy as np
    import torch

    import ensemble as en
    from foam import FoamSim
    cfgs = en.c4_configs(B)
    inputs = [en.c4_inputs(c, b) for b, c in enumerate(cfgs)]
    ens = en.Ensemble(cfgs, inputs)
    tiles = ens.autotune_tiles()
    for _ in range(warmup):
        ens.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ens.step()
    torch.cuda.synchronize()
    el_eager = time.perf_counter() - t0
    ens.capture()                      # the frame as a hip graph
    ens.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ens.replay(check=False)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ens.check()                        # sticky overflow / fixed-point flags
    seq = []
    for b in np.linspace(0, B - 1, sample).astype(int):
        sim = FoamSim(cfgs[b], *inputs[b])
        for _ in range(2):
            sim.step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(3):
            sim.step()
        torch.cuda.synchronize()
        seq.append((time.perf_counter() - t1) / 3)
    ns = np.array([c.n for c in cfgs])
    rs = np.array([c.res for c in cfgs])
    fps = steps / el
    seq_rate = 1.0 / float(np.mean(seq))
    # where the foams' FSC sits against their bands at the end of the run
    mean_fsc = np.array([float(ens.snapshot(b)["fsc"].mean()) for b in range(B)])
    rad_rel = np.array([float(np.mean(ens.snapshot(b)["rad"] / np.float32(cfgs[b].r_max)))
                        for b in range(B)])
    return {"workload": f"{B} foams, N log-uniform [500, 20000], radii U[0.6,1.2]*(0.5/N)^(1/3), "
                        "JFA edge round(5 N^(1/3)) in [48,160], bands {10-14,12-16,14-18}, "
                        "C2 dynamics",
            "foams": B, "particles": int(ns.sum()), "voxels": int((rs ** 3).sum()),
            "frames": steps, "batch_frames_per_s": round(fps, 3),
            "foam_frames_per_s": round(B * fps, 1),
            "eager_foam_frames_per_s": round(B * steps / el_eager, 1),
            "frame": "hip graph replay (eager frames beside it)",
            "jfa_tiles": {"picked": ens.tile_choice, "ms_per_pass_sequence": tiles},
            "one_at_a_time_foam_frames_per_s": round(seq_rate, 1),
            "fsc": {"mean_fsc": [round(float(mean_fsc.min()), 2),
                                 round(float(np.median(mean_fsc)), 2),
                                 round(float(mean_fsc.max()), 2)],
                    "band_low": [int(min(c.f_low for c in cfgs)), int(max(c.f_low for c in cfgs))],
                    "mean_rad_over_r_max": round(float(rad_rel.mean()), 3),
                    "frames_run": int(ens.frame),
                    "note": "the foams' mean FSC stays below their bands (10-18): at a JFA edge "
                            "of 5*N^(1/3) voxels a face spans few voxels against the score "
                            "rule's expected area 1.5*((ri+rj)/voxel)^2 (jfa.py:379-443), so the "
                            "controller keeps asking for growth once the warm-start gate opens "
                            "(frame 30) while r_max = 1.2*(0.5/N)^(1/3) bounds the radii"},
            "speedup": round(B * fps / seq_rate, 2),
            "spread": {"n": [int(ns.min()), int(np.median(ns)), int(ns.max())],
                       "res": [int(rs.min()), int(np.median(rs)), int(rs.max())],
                       "one_at_a_time_ms": [round(1e3 * min(seq), 3),
                                            round(1e3 * float(np.median(seq)), 3),
                                            round(1e3 * max(seq), 3)]}}


def cpu_baseline(cfg, pos, rad, frames, frames_1t=2):
    from oracle import oracle as O
    from oracle.frame import OracleFoam
    threads = len(os.sched_getaffinity(0))
    O.set_threads(threads)
    ref = OracleFoam(cfg, pos, rad)
    t0 = time.perf_counter()
    for _ in range(frames):
        ref.step()
    dt = time.perf_counter() - t0
    out = {"value": round(frames / dt, 4), "unit": "frames/s", "cores": threads,
           "kind": "port", "cpu_model": cpu_model(),
           "sample": f"first {frames} frames of the same N=10k/128^3 workload (JFA every "
                     f"frame), C restatement of the reference loops, OpenMP on {threads} "
                     f"host threads, {dt:.2f} s"}
    if frames_1t > 0:
        O.set_threads(1)
        ref = OracleFoam(cfg, pos, rad)
        t0 = time.perf_counter()
        for _ in range(frames_1t):
            ref.step()
        dt1 = time.perf_counter() - t0
        O.set_threads(threads)
        out["one_thread"] = {"value": round(frames_1t / dt1, 4), "unit": "frames/s",
                             "sample": f"first {frames_1t} frames, 1 thread, {dt1:.2f} s"}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(the foam step is single-device in this build)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import foam_config as fc
    import foam_inputs as fi

    if args.pressure_per_launch:
        import fos_lib as fl
        fl.call("fos_set_param", b"pressure_flow", 0)
    if args.copy_calibration:
        copy_calibration()
    cfg = fc.preset("c2")
    pos, rad = fi.for_config(cfg)
    V = cfg.res ** 3
    P = cfg.passes()
    elapsed, pass_ms, sched_ms, sim = timed(cfg, pos, rad, args.steps, args.warmup, True)
    fps = args.steps / elapsed
    flat = np.array(pass_ms, dtype=np.float64)       # [frames, passes] in ms
    # mean launch time from the frames bracketed as a whole (no record
    # between the passes); the per-pass frames when the run is too short
    avg_pass_s = (float(np.mean(sched_ms)) / P if sched_ms else float(flat.mean())) / 1e3
    achieved = 8.0 * V / avg_pass_s / 1e9            # GB/s
    gvox = P * V / (float(flat.sum(axis=1).mean()) / 1e3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "uniform box N=10000, radii U[0.012,0.024], JFA 128^3 int32, "
                               "8 passes, JFA every frame, PBD<=12, pressure x20, "
                               "Gaussian Brownian 1e-4, OU jitter (run.py:857-867); the frame "
                               "issued by fos_frame_step",
                   "n": cfg.n, "jfa_res": cfg.res, "jfa_passes": P, "hash_res": cfg.hash_res(),
                   "frames": args.steps},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     **traffic_fields(TRAFFIC_FILE), "kernel": "k_jfa_pass_list",
                     "avg_launch_us": round(avg_pass_s * 1e6, 3),
                     "per_pass_us": [round(x * 1e3, 2) for x in flat.mean(axis=0)],
                     "timed_frames_with_events": len(pass_ms),
                     "timed_frames_bracketed": len(sched_ms),
                     **valu_fields(avg_pass_s),
                     "copy_measured_gbs": copy_bandwidth(),
                     "limiter": "instruction issue and latency at full occupancy: VALU 56-70% and "
                                "texture-address 63-65% busy, waves waiting on a dependency 35-42% "
                                "of their cycles (profiles/k_jfa_pass_list_c2_pmc.md)"},
        "gvox_pass_per_s": round(gvox, 2),
        "jfa_ms_per_frame": round(float(flat.sum(axis=1).mean()), 4),
    }
    gate = cfg.warmstart_frames
    out["config"]["window"] = {
        "frames": [args.warmup, args.warmup + args.steps - 1], "growth_gate_frame": gate,
        "frames_past_gate": max(0, args.warmup + args.steps - gate),
        "note": ("the controller (warm-start gate, run.py:1184-1197) is closed before frame "
                 f"{gate}; frames before it run the same kernels except the settle nudge")}
    if not args.no_cadence:
        c2c = fc.preset("c2c")
        # the window starts past the JFA warm start whatever --warmup is
        w2 = max(args.warmup, c2c.jfa_warm)
        el2, _, _, _ = timed(c2c, pos, rad, args.steps, w2, False)
        jfa_frames = sum(1 for f in range(w2, w2 + args.steps) if c2c.jfa_runs(f))
        out["cadence"] = {"workload": "same, JFA every 5th frame after a 20-frame warm start",
                          "value": round(args.steps / el2, 3), "unit": "frames/s",
                          "frames": [w2, w2 + args.steps - 1], "jfa_frames": jfa_frames}
    if not args.no_early_exit:
        out["early_exit"] = early_exit_bench(cfg, pos, rad, args.steps, args.warmup)
    if not args.no_temporal:
        c2t = fc.preset("c2t")
        out["temporal"] = temporal_bench(c2t, pos, rad, args.steps,
                                         max(args.warmup, c2t.skip_warmstart))
    if not args.no_c3:
        out["c3_jfa"] = c3_jfa_bench(args.c3_frames)
        out["c3_frames"] = c3_frames_bench(args.c3_frames)
    if not args.no_surrogate and os.path.exists(WEIGHTS):
        out["surrogate"] = surrogate_bench(args.surrogate_steps, 8)
    if not args.no_ensemble:
        out["ensemble"] = ensemble_bench(args.ensemble_foams, args.ensemble_steps, 3)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg, pos, rad, args.cpu_frames)
        out["speedup_vs_cpu"] = round(fps / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
