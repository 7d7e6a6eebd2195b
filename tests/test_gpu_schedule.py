"""The temporal-skip JFA scheduler (run.py:902-1018) and the adaptive JFA
resolution (run.py:1049-1068) on the GPU against the oracle's frame: the
decision of every frame (run or skip, and why), the resolution of every JFA
frame and the whole state after it.  Integer state bit-exact, positions and
radii within 1e-5 relative."""
import math

import numpy as np
import pytest
import torch

import foam_config as fc
import foam_inputs as fi

pytestmark = pytest.mark.gpu
F = np.float32

INT_STATE = ("face_ids", "face_counts", "fsc", "neighbor_list", "neighbor_count", "used_count")
FLOAT_STATE = ("pos", "rad", "rad_target", "lmd", "fsc_ema", "prev_scores", "v_jit")


def compare(a, b, where):
    for k in INT_STATE:
        assert np.array_equal(a[k], b[k]), (where, k)
    for k in FLOAT_STATE:
        assert np.allclose(a[k], b[k], rtol=1e-5, atol=1e-9), (where, k)


def small_cfg(**kw):
    cfg = fc.preset("c1")
    cfg.n, cfg.res, cfg.warmstart_frames = 600, 40, 1
    cfg.pbd_base, cfg.pbd_max = 4, 8
    cfg.r_min, cfg.r_max = 0.02, 0.07
    cfg.jfa_schedule = "temporal"
    cfg.skip_warmstart, cfg.max_skip = 2, 4
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def test_sched_stats_match_oracle(gpu):
    """fos_sched_stats against the oracle's restatement on random state,
    including particles that crossed the periodic boundary."""
    from oracle.frame import OracleFoam
    n = 5000
    cfg = small_cfg(n=n)
    rng = np.random.default_rng(4)
    pos, rad = fi.uniform(n, 0.01, 0.03, seed=9)
    ref = OracleFoam(cfg, pos, rad)
    ref.pos_prev = (pos + rng.normal(0, 0.002, pos.shape)).astype(F)
    ref.pos_prev[:50, 0] = F(0.4999)
    ref.pos[:50, 0] = F(-0.4998)
    ref.rad_prev = (rad * (1 + rng.normal(0, 0.01, n))).astype(F)
    ref.fsc = rng.integers(0, 30, n).astype(np.int32)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    tp, tr, tf = dev(ref.pos), dev(ref.rad), dev(ref.fsc)
    pp, rp = dev(ref.pos_prev), dev(ref.rad_prev)
    out = torch.zeros(8, dtype=torch.float64, device="cuda")
    ctx = gpu.Context(n, 8, 27)
    for have in (1, 0):
        gpu.call("fos_sched_stats", ctx.handle, gpu.ptr(tp), gpu.ptr(tr), gpu.ptr(tf), n, 1.0,
                 cfg.f_low, cfg.f_high, gpu.ptr(pp), gpu.ptr(rp), have, gpu.ptr(out),
                 gpu.stream())
        got = out.cpu().numpy()
        want = ref.sched_stats(bool(have))
        assert np.allclose(got[:6], want, rtol=1e-12, atol=0), (have, got, want)
        assert got[2] == want[2] and got[5] == want[5]
        # the previous-value buffers now hold the current state
        assert np.array_equal(pp.cpu().numpy(), ref.pos) and np.array_equal(rp.cpu().numpy(), ref.rad)


@pytest.mark.parametrize("path", ["native", "python"])
@pytest.mark.parametrize("variant", ["watchdog", "activity", "mixed"])
def test_temporal_schedule_frames_match_oracle(gpu, path, variant):
    """Frames under the temporal skip: decisions (warm start, skips, the
    watchdog, activity, a band change's lockout) and state equal the
    oracle's.  'watchdog' raises the activity thresholds so only the warm
    start, the watchdog and the band change run the JFA; 'mixed' sets the
    radius-spread threshold to 0.02, about the spread this settling foam
    shows per frame, so frames run on activity or skip (and the out-of-band
    guard gets its turn)."""
    from foam import FoamSim
    from oracle.frame import OracleFoam
    kw = dict(brownian=True, brownian_mode=1, brownian_strength=2e-4, ou_jitter=True)
    if variant == "watchdog":
        kw.update(th_rsigma=10.0, th_margin=10.0, th_fsc_drift=1e9, oob_frac=2.0)
    if variant == "mixed":
        kw.update(th_rsigma=0.02)
    cfg = small_cfg(**kw)
    pos, rad = fi.uniform(cfg.n, 0.03, 0.06, seed=21)
    sim = FoamSim(cfg, pos, rad)
    sim.native = path == "native"
    import copy
    ref = OracleFoam(copy.deepcopy(cfg), pos, rad)
    reasons = []
    for f in range(12):
        if f == 8:
            sim.set_band(cfg.f_low - 1, cfg.f_high + 1)
            ref.set_band(ref.cfg.f_low - 1, ref.cfg.f_high + 1)
        a, b = sim.step(), ref.step()
        print(f"[schedule] {path} {variant} frame={f} jfa={a['jfa']} reason={a['jfa_reason']} "
              f"oracle={b['jfa_reason']} signals={a.get('jfa_signals')}", flush=True)
        assert a["jfa"] == b["jfa"] and a["jfa_reason"] == b["jfa_reason"], f
        assert a["pbd_passes"] == b["pbd_passes"] and a["growth"] == b["growth"], f
        compare(sim.snapshot(), ref.snapshot(), f)
        reasons.append(a["jfa_reason"])
    assert reasons[:2] == ["warm-start"] * 2
    assert reasons[8] == "band-change"
    if variant == "watchdog":
        assert reasons[2:6] == ["skip"] * 3 + ["watchdog"]
    else:
        assert "activity" in reasons
    if variant == "mixed":
        assert "skip" in reasons
    # the reference's decision lines
    lines = sim.log_lines()
    assert lines and lines[0].startswith("[JFA")


def test_adaptive_resolution_frames_match_oracle(gpu):
    """JFA frames at the resolution run.py:1054-1061 picks from the mean
    radius (clamped to [res_min, res_max]), every frame against the oracle."""
    from foam import FoamSim
    from oracle.frame import OracleFoam
    cfg = small_cfg(jfa_schedule="cadence", adaptive_res=True, res_min=24, res_max=48,
                    voxels_per_diameter=3.6, pressure_iters=3)
    pos, rad = fi.uniform(cfg.n, 0.03, 0.06, seed=21)
    sim = FoamSim(cfg, pos, rad)
    ref = OracleFoam(cfg, pos, rad)
    seen = set()
    for f in range(4):
        a, b = sim.step(), ref.step()
        print(f"[adaptive res] frame={f} res={a.get('jfa_res')} oracle={b['jfa_res']}",
              flush=True)
        assert a["jfa"] and a["jfa_res"] == b["jfa_res"], f
        compare(sim.snapshot(), ref.snapshot(), f)
        seen.add(a["jfa_res"])
    r_mean = float(np.mean(rad, dtype=np.float64))
    assert min(seen) >= 24 and max(seen) <= 48
    assert abs(min(seen) - 3.6 / (2 * r_mean)) < 4


def test_temporal_run_equals_steps(gpu):
    """fos_frame_run (frames issued from C++ with deferred face steps) under
    the temporal skip gives the same state as the same number of steps."""
    from foam import FoamSim
    cfg = small_cfg(brownian=True, brownian_mode=1, brownian_strength=2e-4)
    pos, rad = fi.uniform(cfg.n, 0.03, 0.06, seed=5)
    a, b = FoamSim(cfg, pos, rad), FoamSim(cfg, pos, rad)
    for _ in range(9):
        a.step()
    b.run(9)
    sa, sb = a.snapshot(), b.snapshot()
    for k in INT_STATE + FLOAT_STATE + ("sched", "pos_prev", "rad_prev"):
        assert np.array_equal(sa[k], sb[k]), k
