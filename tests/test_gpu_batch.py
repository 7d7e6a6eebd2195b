"""Batched ensembles (no reference counterpart): every foam of a batch gives
what the single-foam path (itself pinned to the oracle) gives on its own
inputs, bit for bit, with batch-global particle ids."""
import numpy as np
import pytest
import torch

import foam_config as fc
import foam_inputs as fi

pytestmark = pytest.mark.gpu


def foams(spec, early=False):
    cfgs, inputs = [], []
    for b, (n, res, rmin, rmax) in enumerate(spec):
        cfg = fc.preset("c1")
        cfg.n, cfg.res, cfg.r_min, cfg.r_max = n, res, rmin, rmax
        cfg.early_exit = early
        cfg.seed = 100 + b
        cfgs.append(cfg)
        inputs.append(fi.uniform(n, rmin, rmax, seed=20 + b))
    return cfgs, inputs


SPEC = [(300, 24, 0.03, 0.06), (700, 40, 0.02, 0.045), (500, 33, 0.025, 0.05),
        (1200, 48, 0.015, 0.035)]


def shift(a, p0):
    """Local particle ids -> batch-global (-1 stays)."""
    a = np.asarray(a).copy()
    a[a >= 0] += p0
    return a


@pytest.mark.parametrize("early,tiles", [(False, "rows"), (True, "rows"), (False, "flat"),
                                         (True, "mixed")])
def test_batch_jfa_matches_single_foams(gpu, early, tiles):
    """Row and flat pass tiles (per foam) give the same labels."""
    from ensemble import Ensemble
    from foam import FoamSim
    cfgs, inputs = foams(SPEC, early)
    print(f"[batch jfa] early={early} tiles={tiles}", flush=True)
    ens = Ensemble(cfgs, inputs)
    if tiles != "rows":
        ens.set_tiles([1] * len(cfgs) if tiles == "flat" else [b % 2 for b in range(len(cfgs))])
    ens.run_jfa()
    torch.cuda.synchronize()
    info = ens.info.cpu().numpy()
    for b, (cfg, (pos, rad)) in enumerate(zip(cfgs, inputs)):
        sim = FoamSim(cfg, pos, rad)
        sim.run_jfa()
        torch.cuda.synchronize()
        p0 = int(ens.p0[b])
        si = sim.info.cpu().numpy()
        ex = int(si[0])
        assert info[b, 0] == ex and info[b, 1] == si[1], b
        assert (info[b, 2:2 + ex] == si[2:2 + ex]).all(), b
        assert np.array_equal(ens.labels(b).cpu().numpy(), shift(sim.grid.cpu().numpy(), p0)), b
        assert np.array_equal(ens.labels(b, ens.filtered).cpu().numpy(),
                              shift(sim.filtered.cpu().numpy(), p0)), b
        for name in ("face_ids", "neighbor_list"):
            got = ens.foam(b, getattr(ens, name)).cpu().numpy()
            assert np.array_equal(got, shift(getattr(sim, name).cpu().numpy(), p0)), (b, name)
        for name in ("face_counts", "fsc", "fsc_ema", "neighbor_count", "used_count",
                     "prev_scores"):
            got = ens.foam(b, getattr(ens, name)).cpu().numpy()
            assert np.array_equal(got, getattr(sim, name).cpu().numpy()), (b, name)
    if early:
        assert len(set(info[:, 0])) > 1 or info[:, 1].any()


def frame_foams():
    """Foams with the C2 dynamics (Gaussian Brownian, adaptive PBD, pressure,
    controller after a short warm start), uneven sizes and bands."""
    spec = [(400, 24, 0.015, 0.045, 10, 14), (900, 36, 0.012, 0.032, 12, 16),
            (650, 30, 0.013, 0.036, 14, 18)]
    cfgs, inputs = [], []
    for b, (n, res, rmin, rmax, lo, hi) in enumerate(spec):
        cfg = fc.preset("c2")
        cfg.ou_jitter = False              # the batch has no OU jitter kernel
        cfg.n, cfg.res, cfg.r_min, cfg.r_max = n, res, rmin, rmax
        cfg.f_low, cfg.f_high = lo, hi
        cfg.warmstart_frames = 1
        cfg.seed = 7 + 13 * b
        cfgs.append(cfg)
        inputs.append(fi.uniform(n, 2 * rmin, rmax, seed=40 + b))
    return cfgs, inputs


@pytest.mark.parametrize("lanes", [0, 8])
def test_batch_frames_match_single_foams(gpu, lanes):
    """Whole frames: every foam of the batch equals its own FoamSim run, bit
    for bit, including foams that run different PBD pass counts.  lanes: the
    PBD pass's lanes per particle (0: by size)."""
    import ctypes
    old = ctypes.c_int()
    gpu.call("fos_get_param", b"pbd_lanes", ctypes.byref(old))
    gpu.call("fos_set_param", b"pbd_lanes", lanes)
    try:
        _batch_frames_match_single_foams()
    finally:
        gpu.call("fos_set_param", b"pbd_lanes", old.value)


def _batch_frames_match_single_foams():
    from ensemble import Ensemble
    from foam import FoamSim
    cfgs, inputs = frame_foams()
    ens = Ensemble(cfgs, inputs)
    sims = [FoamSim(c, p, r) for c, (p, r) in zip(cfgs, inputs)]
    seen = set()
    for f in range(4):
        print(f"[batch frames] frame {f}", flush=True)
        last = ens.step()
        for b, sim in enumerate(sims):
            s = sim.step()
            assert s["pbd_passes"] == last["pbd_passes"][b], (f, b)
            assert s["growth"] == bool(last["growth"][b]), (f, b)
            got, want = ens.snapshot(b), sim.snapshot()
            for k in got:
                assert np.array_equal(got[k], want[k]), (f, b, k)
        seen.add(tuple(last["pbd_passes"]))
    assert any(len(set(p)) > 1 for p in seen) or len(seen) > 1


def test_batch_captured_frames_match_eager(gpu):
    """The frame captured as a hip graph (per-foam decisions on the device,
    fixed finalize rounds) replays to the eager frames bit for bit."""
    from ensemble import Ensemble
    cfgs, inputs = frame_foams()
    a = Ensemble(cfgs, inputs)
    b = Ensemble(cfgs, inputs)
    for _ in range(2):
        a.step()
        b.step()
    b.capture()
    for f in range(4):
        print(f"[batch graph] frame {f}", flush=True)
        a.step()
        b.replay()
        for k in range(len(cfgs)):
            ga, gb = a.snapshot(k), b.snapshot(k)
            for key in ga:
                assert np.array_equal(ga[key], gb[key]), (f, k, key)


def test_batch_tile_autotuner(gpu):
    """The autotuner times its candidates, keeps the fastest, and the frame
    after it equals a batch that kept row tiles."""
    from ensemble import Ensemble
    cfgs, inputs = frame_foams()
    a, b = Ensemble(cfgs, inputs), Ensemble(cfgs, inputs)
    times = b.autotune_tiles(reps=2)
    assert set(times) == {"rows", "flat", "per_edge"} and b.tile_choice in times
    assert all(t > 0 for t in times.values())
    a.step()
    b.step()
    for k in range(len(cfgs)):
        ga, gb = a.snapshot(k), b.snapshot(k)
        for key in ga:
            assert np.array_equal(ga[key], gb[key]), (k, key)


def test_c4_configs_match_single_foams(gpu):
    """The ensemble workload's own foam configurations (N log-uniform in
    [500, 20000], JFA edge round(5 N^(1/3)), per-foam radius bounds, FSC
    bands and seeds; 8 foams): every foam equals its own FoamSim run, bit for
    bit, over three frames."""
    import ensemble as en
    from ensemble import Ensemble
    from foam import FoamSim
    cfgs = en.c4_configs(8, seed=5)
    inputs = [en.c4_inputs(c, b) for b, c in enumerate(cfgs)]
    ens = Ensemble(cfgs, inputs)
    sims = [FoamSim(c, p, r) for c, (p, r) in zip(cfgs, inputs)]
    print("[c4 configs] n =", [c.n for c in cfgs], "res =", [c.res for c in cfgs], flush=True)
    for f in range(3):
        ens.step()
        for b, sim in enumerate(sims):
            sim.step()
            got, want = ens.snapshot(b), sim.snapshot()
            for k in got:
                assert np.array_equal(got[k], want[k]), (f, b, k)
