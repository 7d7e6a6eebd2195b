"""Pins the oracle's temporal-skip JFA scheduler (run.py:902-1018) and adaptive
resolution (run.py:1049-1068) on the CPU with hand-built states whose
decisions and signals follow from the reference's rules by hand."""
import math

import numpy as np

import foam_config as fc
import foam_inputs as fi
from oracle.frame import OracleFoam

F = np.float32


def sim(n=400, **kw):
    cfg = fc.preset("c1")
    cfg.n = n
    cfg.jfa_schedule = "temporal"
    cfg.skip_warmstart, cfg.max_skip = 3, 5
    for k, v in kw.items():
        setattr(cfg, k, v)
    pos, rad = fi.uniform(n, 0.02, 0.04, seed=3)
    o = OracleFoam(cfg, pos, rad)
    o.fsc[:] = (cfg.f_low + cfg.f_high) // 2          # every particle in band
    return o


def decide(o, frames=1):
    out = []
    for _ in range(frames):
        run, res = o.schedule(o.frame)
        out.append((run, o.reason))
        o.frame += 1
    return out


def test_warm_start_then_watchdog_on_a_still_foam():
    """Nothing moves: the first skip_warmstart decisions run (warm start);
    then the JFA is skipped until frames_since_jfa reaches max_skip."""
    o = sim()
    d = decide(o, 3)
    assert d == [(True, "warm-start")] * 3
    d = decide(o, 5)
    # since counts 1..4 (skip), then 5 >= max_skip: watchdog
    assert d == [(False, "skip")] * 4 + [(True, "watchdog")]
    assert o.since == 0


def test_first_decision_forces_the_signals():
    """Without previous values the radius and velocity signals are 999
    (run.py:937-940), so the first decision after the warm start window
    would run on activity alone."""
    o = sim(skip_warmstart=0)
    assert decide(o) == [(True, "activity")]
    # the second decision has previous values and no motion: skip
    assert decide(o) == [(False, "skip")]


def test_radius_spread_signal_is_the_population_std():
    """Half the particles grow by delta: |dr| is 0 or delta, so sigma(|dr|)
    is delta / 2 exactly; the trigger compares sigma / r_mean with
    TH_RSIGMA."""
    o = sim(skip_warmstart=1)
    decide(o)                                # warm start; caches rad
    r_mean = float(o.rad.astype(np.float64).mean())
    delta = F(0.01 * r_mean)                 # sigma/r = ~0.005 > 0.004
    before = o.rad.copy()
    o.rad[::2] += delta
    d = float(np.abs(o.rad[0] - before[0]))  # delta after fp32 rounding
    out = o.sched_stats(True)
    assert math.isclose(out[0], 0.5 * d, rel_tol=1e-3)
    assert math.isclose(out[1], 0.5 * d, rel_tol=1e-3)
    assert (o.rad_prev == o.rad).all()       # the decision cached the new radii
    o2 = sim(skip_warmstart=1)
    decide(o2)
    o2.rad[::2] += F(0.01 * r_mean)
    assert decide(o2) == [(True, "activity")]
    o3 = sim(skip_warmstart=1)
    decide(o3)
    o3.rad[::2] += F(0.006 * r_mean)         # sigma/r = ~0.003 < 0.004
    assert decide(o3) == [(False, "skip")]


def test_displacement_uses_the_minimum_image():
    """A particle that crosses the periodic boundary by a little has a small
    displacement (|d| > L/2 -> d - sign(d) L); one moved by 0.5 r_mean
    triggers (TH_MARGIN_SAFE = 0.4 r_mean)."""
    o = sim(skip_warmstart=1)
    o.pos[0] = (F(0.499), F(0.0), F(0.0))
    decide(o)
    o.pos[0, 0] = F(-0.4995)                 # crossed the boundary by 0.0005
    out = o.sched_stats(True)
    assert math.isclose(out[2], 0.0015, rel_tol=1e-3)
    o = sim(skip_warmstart=1)
    decide(o)
    r_mean = float(o.rad.astype(np.float64).mean())
    o.pos[7, 1] += F(0.5 * r_mean)
    assert decide(o) == [(True, "activity")]
    o = sim(skip_warmstart=1)
    decide(o)
    o.pos[7, 1] += F(0.3 * r_mean)
    assert decide(o) == [(False, "skip")]


def test_fsc_drift_is_zero_after_the_first_jfa():
    """The FSC only changes on JFA frames, so the drift against the mean
    recorded after the last JFA is 0; before any JFA it is 999 (run.py:
    944-949)."""
    o = sim(skip_warmstart=1)
    decide(o)                                 # runs; the next decision records the mean
    o.fsc[:5] += 3
    assert decide(o) == [(False, "skip")]
    assert o.last_fsc_mu == float(o.fsc.astype(np.float64).mean())


def test_out_of_band_guard_needs_a_skipped_frame():
    """More than 10% of particles outside [f_low, f_high] force the JFA once
    the last frame skipped it (frames_since_jfa > 1, run.py:970-977)."""
    o = sim(skip_warmstart=1)
    decide(o)                                 # warm start, since = 0
    o.fsc[: o.n // 5] = 0                     # 20% out of band
    assert decide(o) == [(False, "skip")]    # since = 1: guard not armed
    assert decide(o) == [(True, "out-of-band-guard")]
    o = sim(skip_warmstart=1)
    decide(o)
    o.fsc[: o.n // 20] = 0                    # 5%: below the 10% guard
    assert decide(o, 2) == [(False, "skip")] * 2


def test_band_change_lockout_forces_one_frame():
    """A band change sets the lockout to SKIP_BAND_CHANGE_LOCKOUT = 2; the
    decision decrements it first, so it forces exactly one JFA."""
    o = sim(skip_warmstart=1)
    decide(o, 2)
    o.set_band(o.cfg.f_low - 1, o.cfg.f_high + 1)
    assert decide(o, 2) == [(True, "band-change"), (False, "skip")]


def test_warm_start_outranks_the_other_reasons():
    o = sim(skip_warmstart=2)
    o.set_band(o.cfg.f_low - 1, o.cfg.f_high + 1)
    assert decide(o) == [(True, "warm-start")]


def test_adaptive_resolution_rounds_half_to_even_and_clamps():
    """res = round(L * voxels_per_diameter / (2 r_mean)) with Python's round
    (ties to even), clamped to [res_min, res_max] (run.py:1054-1061)."""
    o = sim(n=64, skip_warmstart=5, adaptive_res=True, res_min=16, res_max=64,
            voxels_per_diameter=12.0)
    o.rad[:] = F(0.25)                        # 12 / 0.5 = 24 exactly
    assert o.schedule(0) == (True, 24)
    o.rad[:] = F(0.125)                       # 48
    assert o.schedule(1) == (True, 48)
    o.rad[:] = F(0.0625)                      # 96 -> clamped to 64
    assert o.schedule(2) == (True, 64)
    o.rad[:] = F(1.0)                         # 6 -> clamped to 16
    assert o.schedule(3) == (True, 16)
    # a tie: 12 / (2 r) = 24.5 rounds to 24 (even), 25.5 to 26
    o.cfg.voxels_per_diameter = 12.25
    o.rad[:] = F(0.25)
    assert o.schedule(4) == (True, 24)
    o.cfg.voxels_per_diameter = 12.75
    o.cfg.skip_warmstart = 10
    assert o.schedule(5) == (True, 26)


def test_cadence_schedule_is_unchanged():
    o = sim(jfa_schedule="cadence", jfa_every=3, jfa_warm=2)
    got = [o.schedule(f)[0] for f in range(8)]
    assert got == [True, True, True, False, False, True, False, False]
