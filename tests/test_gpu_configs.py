"""The benchmark workloads at their own sizes against the oracle's frame.

C3 (N=200,000 clustered, 384^3), C5 (N=50,000, 256^3) and the C4 batched
ensemble: whole frames in the run.py:700-1261 order, every state array
compared after each frame -- integer state (face rows, FSC, neighbour lists)
bit-exact, positions and radii within 1e-5 relative.  Also the velocity
paths (XSPH, damping: dynamics.py:473-480, 532-593) and the radius range of
compute_pressure_stats (dynamics.py:1561-1589) against their definitions.
"""
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
    """Every state array both sides hold (a batch keeps no OU velocities)."""
    for k in INT_STATE:
        assert np.array_equal(a[k], b[k]), (where, k, int((a[k] != b[k]).sum()))
    for k in FLOAT_STATE:
        if k in a and k in b:
            assert np.allclose(a[k], b[k], rtol=1e-5, atol=1e-9), (where, k)


def frames_against_oracle(cfg, pos, rad, frames, tag):
    from foam import FoamSim
    from oracle import oracle as O
    from oracle.frame import OracleFoam
    import os
    O.set_threads(min(16, len(os.sched_getaffinity(0))))
    sim = FoamSim(cfg, pos, rad)
    ref = OracleFoam(cfg, pos, rad)
    for f in range(frames):
        print(f"[{tag}] frame {f}", flush=True)
        a, b = sim.step(), ref.step()
        assert a["pbd_passes"] == b["pbd_passes"] and a["jfa"] == b["jfa"], f
        assert a["growth"] == b["growth"], f
        sa, sb = sim.snapshot(), ref.snapshot()
        compare(sa, sb, f)
        assert sa["settle"] == sb["settle"], f
        print(f"[{tag}] frame {f} ok: pbd={a['pbd_passes']} mean FSC {sa['fsc'].mean():.2f}",
              flush=True)
    return sim


@pytest.mark.timeout(900)
def test_c3_frames_match_oracle(gpu):
    """The large clustered foam: grid rebuilds over crowded hash cells, PBD,
    max overlap, pressure, 10-pass JFA with ids above 65,535, face rows and
    finalize, two whole frames."""
    cfg = fc.preset("c3")
    pos, rad = fi.for_config(cfg)
    frames_against_oracle(cfg, pos, rad, 2, "c3 frames")


@pytest.mark.timeout(600)
def test_c5_frames_match_oracle(gpu):
    """The predictor workload's plain frames (N=50,000, 256^3, 20 pressure
    iterations, Gaussian Brownian noise)."""
    cfg = fc.preset("c5")
    pos, rad = fi.for_config(cfg)
    frames_against_oracle(cfg, pos, rad, 2, "c5 frames")


@pytest.mark.timeout(600)
def test_c5_predictor_frames_full_size(gpu):
    """The predictor path at its own size with the committed weights: a judge
    frame (JFA, then the predictor beside it) whose JFA state equals the
    oracle's, then predicted frames whose FSC and face rows satisfy what the
    predictor defines: a kept pair counts for both particles, rows list
    partners j > i once each with count 1, and FSC stays close to the JFA's
    on the next judge frame."""
    import json
    import os
    from foam import FoamSim
    from oracle.frame import OracleFoam
    w = np.load(os.path.join(os.path.dirname(fc.__file__), "face_mlp_c5.npz"), allow_pickle=False)
    params, meta = w["params"], json.loads(str(w["meta"]))
    cfg = fc.preset("c5s")
    pos, rad = fi.for_config(cfg)
    sim = FoamSim(cfg, pos, rad)
    sim.set_predictor(params, meta["mult"])
    ref = OracleFoam(fc.preset("c5"), pos, rad)
    print("[c5 predictor] judge frame 0", flush=True)
    a, b = sim.step(), ref.step()
    assert a["jfa"] and b["jfa"] and a["pbd_passes"] == b["pbd_passes"]
    compare(sim.snapshot(), ref.snapshot(), "judge frame")
    for f in range(1, cfg.judge_every + 1):
        print(f"[c5 predictor] frame {f}", flush=True)
        last = sim.step()
        torch.cuda.synchronize()
        if last["jfa"]:
            break
        fsc = sim.fsc.cpu().numpy().astype(np.int64)
        ids = sim.face_ids.cpu().numpy()
        cts = sim.face_counts.cpu().numpy()
        n = cfg.n
        kept = ids >= 0
        rows = np.repeat(np.arange(n)[:, None], 32, axis=1)
        assert (ids[kept] > rows[kept]).all(), f          # partners j > i
        assert (cts[kept] == 1).all() and (cts[~kept] == 0).all(), f
        # each row holds its kept partners first, then -1
        assert (np.diff(kept.astype(np.int8), axis=1) <= 0).all(), f
        for i in range(0, n, 997):
            r = ids[i][kept[i]]
            assert len(set(r.tolist())) == len(r), (f, i)
        # FSC counts a kept pair at both ends: the partners a row lists plus
        # the rows that list the particle, exactly unless a row was cut at 32
        listed = kept.sum(axis=1)
        incoming = np.bincount(ids[kept], minlength=n)
        assert (fsc >= listed + incoming).all(), f
        if (listed < 32).all():
            assert np.array_equal(fsc, listed + incoming), f
    jf, err, bias = sim.judged[-1]
    assert jf == cfg.judge_every, sim.judged
    print(f"[c5 predictor] judge frame {jf}: per-particle {err:.4f}, mean FSC {bias:.4f}",
          flush=True)
    assert err < 0.10 and bias < 0.03, (err, bias)


def _ensemble_vs_oracle(cfgs, inputs, frames, sample, tag):
    from ensemble import Ensemble
    from oracle import oracle as O
    from oracle.frame import OracleFoam
    import os
    O.set_threads(min(16, len(os.sched_getaffinity(0))))
    ens = Ensemble(cfgs, inputs)
    refs = {b: OracleFoam(cfgs[b], *inputs[b]) for b in sample}
    for f in range(frames):
        print(f"[{tag}] frame {f}", flush=True)
        ens.step()
        for b, ref in refs.items():
            ref.step()
            compare(ens.snapshot(b), ref.snapshot(), (f, b))
    return ens


@pytest.mark.timeout(600)
def test_c4_configs_match_oracle(gpu):
    """Eight foams of the ensemble workload (N log-uniform in [500, 20000],
    JFA edge round(5 N^(1/3)), their own radius bounds, FSC bands and seeds)
    advanced as one batch, every foam against its own oracle frame."""
    import ensemble as en
    cfgs = en.c4_configs(8, seed=5)
    inputs = [en.c4_inputs(c, b) for b, c in enumerate(cfgs)]
    print("[c4 oracle] n =", [c.n for c in cfgs], "res =", [c.res for c in cfgs], flush=True)
    _ensemble_vs_oracle(cfgs, inputs, 3, range(len(cfgs)), "c4 oracle")


@pytest.mark.timeout(600)
def test_c4_full_batch_sample_matches_oracle(gpu):
    """The full 256-foam batch for two frames; the smallest, largest and a
    spread of foams between them against their oracle frames."""
    import ensemble as en
    cfgs = en.c4_configs(256)
    inputs = [en.c4_inputs(c, b) for b, c in enumerate(cfgs)]
    order = np.argsort([c.n for c in cfgs])
    sample = sorted({int(order[k]) for k in (0, 51, 102, 153, 204, 255)} | {0, 255})
    print("[c4 batch] sample", sample, [cfgs[b].n for b in sample], flush=True)
    _ensemble_vs_oracle(cfgs, inputs, 2, sample, "c4 batch")


def test_xsph_damping_and_radius_range(gpu, orc, monkeypatch):
    """XSPH smoothing over a non-zero velocity field (neighbours within
    2 r_max in the 27 hash cells, blended by eps) bit-exact against the
    oracle, damping as the fp32 product, and the radius range of
    compute_pressure_stats as numpy's min and max."""
    import dynamics
    fl = gpu
    n, G_, L = 3000, 14, 1.0
    pos, rad = fi.uniform(n, 0.02, 0.035, seed=44)
    p, cc, cs, ci = orc.grid_build(pos, n, G_, L)
    rng = np.random.default_rng(3)
    vel = rng.normal(0, 1e-3, (n, 3)).astype(F)
    vel[::7] = 0.0                                  # some particles at rest
    lim = fc.f32((2.0 * 0.035) ** 2)
    want = orc.xsph(p, vel, n, G_, L, lim, fc.XSPH_EPSILON, cs, cc, ci)
    t = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda()
         for k, v in dict(p=p, v=vel, cs=cs, cc=cc, ci=ci).items()}
    vt = torch.zeros_like(t["v"])
    fl.call("fos_xsph", fl.ptr(t["p"]), fl.ptr(t["v"]), fl.ptr(vt), n, G_, L, lim,
            fc.XSPH_EPSILON, fl.ptr(t["cs"]), fl.ptr(t["cc"]), fl.ptr(t["ci"]), fl.stream())
    got = t["v"].cpu().numpy()
    assert np.array_equal(got, want)
    assert not np.array_equal(got, vel)             # the smoothing did something
    # the drop-in (R_MAX sets the neighbour radius) gives the same field
    monkeypatch.setattr(fc, "R_MAX", 0.035)
    monkeypatch.setattr(fc, "GRID_RES", G_)
    monkeypatch.setattr(fc, "DOMAIN_SIZE", L)
    v2 = torch.from_numpy(vel).cuda()
    dynamics.apply_xsph_smoothing(t["p"], v2, torch.zeros_like(v2), None, t["cs"], t["cc"],
                                  t["ci"], n)
    assert np.array_equal(v2.cpu().numpy(), want)
    # damping (dynamics.py:473-480): vel *= GLOBAL_DAMPING
    dynamics.apply_global_damping(v2, n)
    assert np.array_equal(v2.cpu().numpy(), (want * F(fc.GLOBAL_DAMPING)).astype(F))
    # zero velocities stay zero through both
    z = torch.zeros_like(v2)
    fl.call("fos_xsph", fl.ptr(t["p"]), fl.ptr(z), fl.ptr(vt), n, G_, L, lim,
            fc.XSPH_EPSILON, fl.ptr(t["cs"]), fl.ptr(t["cc"]), fl.ptr(t["ci"]), fl.stream())
    assert not z.cpu().numpy().any()
    # compute_pressure_stats: (min radius, max radius) over the first n
    r = rng.uniform(0.004, 0.02, 70001).astype(F)
    r[12345], r[54321] = F(0.0031), F(0.0277)
    tr = torch.from_numpy(r).cuda()
    lo, hi = dynamics.compute_pressure_stats(70001, tr, 3.0, 0.002, 0.03)
    assert (lo, hi) == (float(r.min()), float(r.max())) == (float(F(0.0031)), float(F(0.0277)))
    lo, hi = dynamics.compute_pressure_stats(1000, tr, 3.0, 0.002, 0.03)
    assert (lo, hi) == (float(r[:1000].min()), float(r[:1000].max()))
