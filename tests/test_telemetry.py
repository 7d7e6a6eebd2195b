"""Telemetry line formats against the reference's f-strings (run.py:1095-1163,
1303-1318), written out by hand for fixed inputs."""
import numpy as np

import telemetry as tm


STATS = {"mean_fsc": 13.456, "min_fsc": 4, "max_fsc": 22, "mean_fsc_ema": 13.04,
         "mean_score": 1.0349, "std_score": 0.4151, "overflow_pct": 0.0, "p_dir": 67280,
         "p_und_accepted": 33640, "num_passes": 7, "early_exit": True,
         "delta_history": [0.91, 0.5, 0.2012, 0.1, 0.05, 0.012, 0.0004]}
VALID = {"passed": True, "asymmetric_pairs": 12, "overflow_count": 0, "self_loops": 0}


def test_jfa_step4_lines():
    got = tm.jfa_step4_lines(STATS, VALID, 0.00123, 128, 10000)
    assert got == [
        "[JFA Step 4] FSC μ=13.5 [4,22] | EMA μ=13.0",
        "      score: μ=1.03 σ=0.42 | pairs: 33640 (P_dir=67280)",
        "      asym=0.1% overflow=0.0% | time=1.2ms | res=128³ | passes=7/8 (early_exit=True)"
        " | δ=[91.0%, 50.0%, 20.1%, 10.0%, 5.0%, 1.2%]",
        "      ✅ Step 4 SUCCESS - All metrics in target range",
    ]
    assert tm.parse_step4(got)["p_dir"] == 67280


def test_jfa_step4_status_and_deltas():
    s = dict(STATS, early_exit=False, delta_history=[0.25, 0.125], mean_score=1.5)
    got = tm.jfa_step4_lines(s, VALID, 0.0, 100, 10000)
    assert got[2].endswith("passes=7/8 (early_exit=False) | δ=[25.0%, 12.5%]")
    assert got[3] == "      ✓ Validation passed (still converging)"
    got = tm.jfa_step4_lines(dict(s, delta_history=[]), dict(VALID, passed=False), 0.0, 256,
                             10000)
    assert got[2].endswith("passes=7/9 (early_exit=False)")
    assert got[3] == "      ⚠️  Validation FAILED"
    # a single early-exit delta gets the star
    got = tm.jfa_step4_lines(dict(STATS, delta_history=[0.005]), VALID, 0.0, 64, 1)
    assert got[2].endswith("| δ=[0.5%*]") and "asym=1200.0%" in got[2]


def test_dirty_tiles_and_frame_lines():
    assert tm.dirty_tiles_line(37, 512, 1, 0.25, 0.1, 128) == (
        "[JFA][PhaseA] tiles=37/512 (7.2%) | halo=1 | Δpos_thr=0.25×0.007812 | "
        "Δr_thr=0.1×0.007812")
    fsc = np.array([10, 12, 17], np.int32)
    rad = np.array([0.01, 0.012, 0.02], np.float32)
    assert tm.frame_lines(100, 3, 6, 0.0012345, fsc, rad, 12, 16) == [
        "[Frame  100] N=3, Passes=6, MaxDepth=0.001234",
        "    FSC: mean=13.0, min=10, max=17 | band=[12,16]",
        "    Radius: mean=0.0140, min=0.0100, max=0.0200",
        "    Rescue: 0.0% (disabled)",
    ]
    assert tm.frame_lines(1200, 3, 6, 0.0, fsc, rad, 12, 16, 2.5, True)[3] == "    Rescue: 2.5%"
