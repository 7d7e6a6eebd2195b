"""Whole-frame parity (run.py:700-1261 order) and the drop-in modules.

The GPU driver and the CPU restatement advance the same state; after each
frame every state array is compared: integer arrays bit-exact, positions and
radii within 1e-5 relative (and, in practice, bit-exact).
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


@pytest.fixture
def keep_config():
    """Restores the foam_config constants a drop-in test rebinds."""
    saved = {k: v for k, v in vars(fc).items() if k.isupper()}
    yield
    for k, v in saved.items():
        setattr(fc, k, v)


def compare(a, b, where):
    for k in INT_STATE:
        assert np.array_equal(a[k], b[k]), (where, k)
    for k in FLOAT_STATE:
        assert np.allclose(a[k], b[k], rtol=1e-5, atol=1e-9), (where, k)


def small_cfg(**kw):
    cfg = fc.preset("c1")
    cfg.n = 600
    cfg.res = 40
    cfg.frames = 4
    cfg.warmstart_frames = 1
    cfg.pbd_base, cfg.pbd_max = 4, 8
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


@pytest.mark.parametrize("variant", ["plain", "noise", "cadence", "global_back", "early_exit"])
def test_frames_match_oracle(gpu, variant):
    """Whole frames against the oracle's.  'early_exit' runs the reference's
    default early exit (config.py:208-210) with a threshold of 3% so that it
    fires on this small foam: the passes each frame ran and their change
    history must match too."""
    from foam import FoamSim
    from oracle.frame import OracleFoam
    kw = {}
    if variant == "early_exit":
        kw = dict(early_exit=True, th_pass_delta=0.03, min_passes=3)
    if variant == "noise":
        kw = dict(brownian=True, brownian_mode=1, brownian_strength=1e-3, ou_jitter=True)
    if variant == "cadence":
        kw = dict(jfa_every=2, jfa_warm=1, pressure_iters=3)
    if variant == "global_back":          # settle with run.py's global backpressure
        kw = dict(backpressure_mode=1)
    cfg = small_cfg(**kw)
    pos, rad = fi.uniform(cfg.n, 0.03, 0.06, seed=21)
    cfg.r_min, cfg.r_max = 0.02, 0.07
    sim = FoamSim(cfg, pos, rad)
    ref = OracleFoam(cfg, pos, rad)
    ran = []
    for f in range(cfg.frames):
        print(f"[frame parity] variant={variant} frame={f}", flush=True)
        a = sim.step()
        b = ref.step()
        assert a["pbd_passes"] == b["pbd_passes"] and a["jfa"] == b["jfa"]
        assert a["growth"] == b["growth"]
        sa, sb = sim.snapshot(), ref.snapshot()
        compare(sa, sb, f)
        assert sa["settle"] == sb["settle"], f
        print(f"[frame parity] variant={variant} frame={f} settle={int(sa['settle'])}", flush=True)
        if a["jfa"]:
            info = sim.info.cpu().numpy()
            passes, early, hist = ref.jfa_info
            assert int(info[0]) == passes and bool(info[1]) == bool(early), f
            assert [int(x) for x in info[2:2 + passes]] == [int(x) for x in hist[:passes]], f
            ran.append(passes)
    if variant == "early_exit":
        assert min(ran) < cfg.passes(), ran


def test_c1_first_frames_bit_exact(gpu):
    from foam import FoamSim
    from oracle.frame import OracleFoam
    cfg = fc.preset("c1")
    pos, rad = fi.for_config(cfg)
    sim = FoamSim(cfg, pos, rad)
    ref = OracleFoam(cfg, pos, rad)
    for f in range(2):
        sim.step()
        ref.step()
        a, b = sim.snapshot(), ref.snapshot()
        compare(a, b, f)
        assert np.array_equal(a["pos"], b["pos"]) and np.array_equal(a["fsc"], b["fsc"])


def test_c2_frames_match_oracle(gpu):
    """The benchmark workload itself (N=10,000, 128^3 JFA every frame, 20
    pressure iterations, Brownian noise): three whole frames against the
    oracle's frame, integer state bit-exact."""
    from foam import FoamSim
    from oracle.frame import OracleFoam
    cfg = fc.preset("c2")
    pos, rad = fi.for_config(cfg)
    sim = FoamSim(cfg, pos, rad)
    ref = OracleFoam(cfg, pos, rad)
    for f in range(3):
        print(f"[c2 frames] frame {f}", flush=True)
        a, b = sim.step(), ref.step()
        assert a["pbd_passes"] == b["pbd_passes"] and a["growth"] == b["growth"], f
        compare(sim.snapshot(), ref.snapshot(), f)


def test_frame_start_reuses_the_end_of_frame_grid(gpu):
    """A frame starts from the CSR the last frame's final rebuild left while
    nothing has written the positions since: that CSR equals a fresh rebuild,
    and a torch store into the positions makes the next frame rebuild."""
    from foam import FoamSim
    cfg = small_cfg()
    pos, rad = fi.uniform(cfg.n, 0.03, 0.06, seed=11)
    sim = FoamSim(cfg, pos, rad)
    sim.step()
    assert sim.grid_key == sim.grid_token()
    kept = [t.clone() for t in (sim.cell_count, sim.cell_start, sim.cell_indices, sim.pos)]
    sim.rebuild_grid()
    fresh = (sim.cell_count, sim.cell_start, sim.cell_indices, sim.pos)
    for a, b in zip(kept, fresh):
        assert torch.equal(a, b)
    sim.pos[0, 0] += 0.25
    assert sim.grid_key != sim.grid_token()
    # a copy with the same store and a forced rebuild steps identically
    twin = FoamSim(cfg, pos, rad)
    twin.restore(dict(sim.snapshot()))
    sim.step()
    twin.step()
    compare(sim.snapshot(), twin.snapshot(), "after a store")


def test_snapshot_roundtrip(gpu, tmp_path):
    from foam import FoamSim
    cfg = small_cfg()
    pos, rad = fi.uniform(cfg.n, 0.03, 0.06, seed=5)
    a = FoamSim(cfg, pos, rad)
    a.step()
    path = str(tmp_path / "s.npz")
    a.save(path)
    b = FoamSim(cfg, pos, rad)
    b.load(path)
    a.step()
    b.step()
    compare(a.snapshot(), b.snapshot(), "restore")
    a.export_csv(str(tmp_path / "p.csv"))
    head = open(tmp_path / "p.csv").readline().strip()
    assert head == "ID,X,Y,Z,Radius,FSC"


def test_telemetry_lines(gpu):
    """Reference-format log lines of a frame: the JFA block parses back to
    the frame's FSC statistics and the pass count of the JFA that ran."""
    import telemetry as tm
    from foam import FoamSim
    cfg = small_cfg()
    pos, rad = fi.uniform(cfg.n, 0.03, 0.06, seed=6)
    sim = FoamSim(cfg, pos, rad, telemetry=True)
    last = sim.step()
    lines = sim.log_lines()
    assert lines[0].startswith("[JFA Step 4] FSC")
    got = tm.parse_step4(lines[:3])
    fsc = sim.fsc.cpu().numpy()
    assert got["min_fsc"] == fsc.min() and got["max_fsc"] == fsc.max()
    assert abs(got["mean_fsc"] - fsc.mean()) <= 0.05
    assert got["num_passes"] == last["num_passes"] and got["p_dir"] == last["p_dir"]
    assert len(last["delta_history"]) == last["num_passes"]
    sim.frame = 100                      # the every-100-frames block
    assert sim.log_lines()[4].startswith("[Frame  100] N=")


def test_dropin_run_jfa_matches_oracle(gpu, orc, keep_config):
    import jfa
    fc.JFA_RES_MIN, fc.JFA_RES_MAX = 16, 320
    fc.DOMAIN_SIZE = 1.0
    fc.JFA_EARLY_EXIT_ENABLED = False
    n, res = 2000, 64
    pos, rad = fi.uniform(n, 0.02, 0.04, seed=0)
    jfa.init_jfa(max_particles=n, max_res=res)
    jfa.set_resolution(res)
    assert jfa.JFA_NUM_PASSES == 7
    tp = torch.from_numpy(pos).cuda()
    tr = torch.from_numpy(rad).cuda()
    fsc, stats = jfa.run_jfa(tp, tr, n)
    voxel = fc.f32(1 / 64)
    lab, passes, _, hist = orc.jfa_run(pos, rad, n, res, 1.0, voxel, 7)
    ids, cts = orc.accumulate_faces(orc.majority(lab, res), res, n)
    prev, nl, nc, used, fs = orc.finalize_faces(n, rad, voxel, ids, cts, np.zeros((n, 32), F))
    assert np.array_equal(jfa.jfa_grid.to_numpy().ravel(), lab)
    assert np.array_equal(fsc, fs) and np.array_equal(jfa.face_ids.to_numpy(), ids)
    assert np.array_equal(jfa.neighbor_list.to_numpy(), nl)
    assert stats["num_passes"] == 7 and not stats["early_exit"]
    assert stats["delta_history"] == [h / res ** 3 for h in hist]
    assert stats["min_fsc"] == fs.min() and stats["max_fsc"] == fs.max()
    v = jfa.validate_jfa(n)
    assert v["self_loops"] == 0 and v["asymmetric_pairs"] == int(orc.validate(n, nl, nc)[0])
    # the per-kernel entry points reproduce the same grid
    jfa.jfa_init(tp, tr, n)
    step = res // 2
    for p in range(7):
        (jfa.jfa_step_ping0 if p % 2 == 0 else jfa.jfa_step_ping1)(tp, tr, max(step, 1))
        step //= 2
    jfa.jfa_copy_temp_to_grid()
    assert np.array_equal(jfa.jfa_grid.to_numpy().ravel(), lab)
    assert jfa.label_changes[None] == hist[-1]


def test_dropin_grid_and_dynamics(gpu, orc, keep_config):
    import dynamics
    import grid
    n, G_ = 2000, 13
    fc.DOMAIN_SIZE, fc.GRID_RES = 1.0, G_
    pos, rad = fi.uniform(n, 0.02, 0.04, seed=0)
    p, cc, cs, ci = orc.grid_build(pos, n, G_, 1.0)
    tp = torch.from_numpy(pos).cuda()
    tr = torch.from_numpy(rad).cuda()
    tcc = torch.zeros(G_ ** 3, dtype=torch.int32, device="cuda")
    tcs = torch.zeros_like(tcc)
    tcw = torch.zeros_like(tcc)
    tci = torch.zeros(n, dtype=torch.int32, device="cuda")
    grid.clear_grid(tcc)
    grid.count_particles_per_cell(tp, tcc, n)
    grid.prefix_sum(tcc, tcs)
    grid.copy_cell_pointers(tcs, tcw)
    grid.scatter_particles(tp, tcw, tci, n)
    assert np.array_equal(tci.cpu().numpy(), ci) and np.array_equal(tp.cpu().numpy(), p)
    lmd = torch.zeros(n, device="cuda")
    g, want_l = orc.max_overlap(p, rad, n, G_, 1.0, cs, cc, ci)
    assert dynamics.compute_max_overlap(tp, tr, tcs, tcc, tci, lmd, n) == g
    dynamics.project_overlaps(tp, tr, tcs, tcc, tci, n)
    want = orc.pbd_jacobi(p, rad, n, G_, 1.0, fc.f32(1.015), 0.2, 2.0, cs, cc, ci)
    assert np.array_equal(tp.cpu().numpy(), want)
    assert dynamics.hash_func(7, 3) == orc.lcg_hash(7, 3)


def test_runs_are_deterministic(gpu):
    """Kernels are deterministic by construction (atomics only on integers
    or order-free maxima, fp32 sums in fixed orders): two fresh runs of the
    same frames agree bit for bit in every state array, floats included --
    the plain C2 frame, the learned-predictor frame path and a batch."""
    import face_predictor as fp
    from ensemble import Ensemble
    from foam import FoamSim

    def plain():
        cfg = fc.preset("c2")
        pos, rad = fi.for_config(cfg)
        sim = FoamSim(cfg, pos, rad)
        for _ in range(4):
            sim.step()
        return [sim.snapshot()]

    def predicted():
        cfg = fc.preset("c5s")
        cfg.n, cfg.res, cfg.judge_every, cfg.pressure_iters = 3000, 64, 3, 5
        cfg.r_min, cfg.r_max = 0.012, 0.03
        pos, rad = fi.uniform(cfg.n, 0.015, 0.025, seed=9)
        sim = FoamSim(cfg, pos, rad)
        sim.set_predictor(fp.init_params(seed=3), mult=2.0)
        for _ in range(5):
            sim.step()
        return [sim.snapshot()]

    def batch():
        cfgs, inputs = [], []
        for b, (n, res) in enumerate(((500, 40), (1500, 48), (900, 32))):
            cfg = small_cfg()
            cfg.n, cfg.res, cfg.seed = n, res, 5 + b
            cfgs.append(cfg)
            inputs.append(fi.uniform(n, 0.03, 0.06, seed=60 + b))
        ens = Ensemble(cfgs, inputs)
        for _ in range(3):
            ens.step()
        return [ens.snapshot(b) for b in range(len(cfgs))]

    for run in (plain, predicted, batch):
        print(f"[determinism] {run.__name__}", flush=True)
        a, b = run(), run()
        for sa, sb in zip(a, b):
            for k in sa:
                assert np.array_equal(sa[k], sb[k]), (run.__name__, k)


def shell_foam(n=300, big=0.15, small=0.012):
    """One large particle at the centre and n - 1 small ones on a shell that
    touches it (Fibonacci points): row 0 meets far more face partners than
    a fresh context's 64 row-table slots."""
    k = np.arange(n - 1, dtype=np.float64) + 0.5
    phi = np.arccos(1.0 - 2.0 * k / (n - 1))
    th = np.pi * (1.0 + 5.0 ** 0.5) * k
    u = np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)
    pos = np.concatenate([np.zeros((1, 3)), (big + small) * u]).astype(F)
    rad = np.full(n, small, F)
    rad[0] = big
    return fi.wrap(pos), rad


@pytest.mark.parametrize("variant", ["c2", "noise_cadence", "short_burst", "table_growth"])
def test_frame_run_equals_steps(gpu, variant):
    """FoamSim.run(k) (one fos_frame_run call, each frame's face step left
    to the next frame to finish) against k step() calls: every state array
    and the last record bit for bit, across runs of several lengths with the
    warm-start gate opening inside one, then step() after a run.
    short_burst: two finalize rounds before the flags are read, so the
    deferred steps finish their rounds on the host; table_growth: a fresh
    context's face-row tables overflow on the first frame, which a deferred
    step redoes after growing them."""
    from foam import FoamSim
    if variant == "c2":
        cfg = fc.preset("c2")
        cfg.warmstart_frames = 4
        pos, rad = fi.for_config(cfg)
    elif variant == "table_growth":
        cfg = small_cfg(n=300, res=64)
        cfg.r_min, cfg.r_max = 0.005, 0.2
        pos, rad = shell_foam()
    else:
        cfg = small_cfg(brownian=True, brownian_mode=1, brownian_strength=1e-3, ou_jitter=True,
                        jfa_every=3, jfa_warm=2)
        cfg.r_min, cfg.r_max = 0.02, 0.07
        pos, rad = fi.uniform(cfg.n, 0.03, 0.06, seed=34)
    if variant == "short_burst":
        gpu.call("fos_set_param", b"fin_first_burst", 2)
    try:
        a, b = FoamSim(cfg, pos, rad), FoamSim(cfg, pos, rad)
        run_and_compare(a, b, variant)
    finally:
        gpu.call("fos_set_param", b"fin_first_burst", 0)
    if variant == "table_growth":
        # the first frame's grid gives particle 0 more than 64 distinct face
        # partners, the row-table slots of a fresh context
        c = FoamSim(cfg, pos, rad)
        c.step()
        g = c.filtered.cpu().numpy().reshape(cfg.res, cfg.res, cfg.res)
        partners = set()
        for ax in range(3):
            h = np.roll(g, -1, axis=ax)
            partners |= set(h[g == 0].tolist()) | set(g[h == 0].tolist())
        partners.discard(0)
        partners.discard(-1)
        assert len(partners) > 64, len(partners)


def run_and_compare(a, b, variant):
    for k in (3, 1, 5, 2):
        print(f"[frame run] {variant} run of {k} from frame {a.frame}", flush=True)
        ra = a.run(k)
        for _ in range(k):
            rb = b.step()
        for key in ("frame", "pbd_passes", "max_depth", "jfa", "growth"):
            assert ra[key] == rb[key], (k, key, ra[key], rb[key])
        sa, sb = a.snapshot(), b.snapshot()
        for key in sa:
            assert np.array_equal(sa[key], sb[key]), (k, key)
    ra, rb = a.step(), b.step()
    assert ra["growth"] == rb["growth"] and ra["pbd_passes"] == rb["pbd_passes"]
    sa, sb = a.snapshot(), b.snapshot()
    for key in sa:
        assert np.array_equal(sa[key], sb[key]), key


def test_integration_snippet_runs(gpu, orc, keep_config):
    """The drop-in snippet of INTEGRATION.md section 2, executed as written
    (REPO set to this checkout): its run_jfa FSC equals the oracle's on the
    same inputs, and the grid, PBD and pressure calls that follow run."""
    import os
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    text = open(os.path.join(root, "INTEGRATION.md")).read()
    sec = text[text.index("## 2. Drop-in for `run.py`"):text.index("## 3.")]
    code = re.search(r"```python\n(.*?)```", sec, re.S).group(1)
    assert 'REPO = "/path/to/repo"' in code
    code = code.replace('REPO = "/path/to/repo"', f"REPO = {root!r}")
    ns = {}
    exec(compile(code, "INTEGRATION.md section 2", "exec"), ns)
    N, RES = ns["N"], ns["RES"]
    assert (N, RES) == (10000, 128) and ns["stats"]["num_passes"] == 8
    voxel = fc.f32(1.0 / RES)
    lab, _, _, _ = orc.jfa_run(ns["pos_np"], ns["rad_np"], N, RES, 1.0, voxel, 8)
    ids, cts = orc.accumulate_faces(orc.majority(lab, RES), RES, N)
    _, _, _, _, fs = orc.finalize_faces(N, ns["rad_np"], voxel, ids, cts, np.zeros((N, 32), F))
    assert np.array_equal(ns["fsc"], fs)
    assert ns["max_dr"] >= 0.0 and ns["changed"] >= 0


@pytest.mark.parametrize("variant", ["c2", "noise_cadence", "global_back", "no_pressure"])
def test_native_frame_equals_python_frame(gpu, variant):
    """fos_frame_step (the frame issued from C++) against the same frame
    issued call by call from Python: every state array and every frame's
    record (PBD passes, overlap maximum, JFA, gate) bit for bit, across a
    snapshot/restore, with the warm-start gate opening inside the run."""
    from foam import FoamSim
    if variant == "c2":
        cfg = fc.preset("c2")
        cfg.warmstart_frames = 3
        pos, rad = fi.for_config(cfg)
    else:
        kw = {"noise_cadence": dict(brownian=True, brownian_mode=1, brownian_strength=1e-3,
                                    ou_jitter=True, jfa_every=2, jfa_warm=2),
              "global_back": dict(backpressure_mode=1),
              "no_pressure": dict(pressure=False, pbd_base=0, pbd_max=0)}[variant]
        cfg = small_cfg(**kw)
        cfg.r_min, cfg.r_max = 0.02, 0.07
        pos, rad = fi.uniform(cfg.n, 0.03, 0.06, seed=33)
    a, b = FoamSim(cfg, pos, rad), FoamSim(cfg, pos, rad)
    b.native = False
    for f in range(7):
        print(f"[native frame] {variant} frame {f}", flush=True)
        ra, rb = a.step(), b.step()
        for k in ("pbd_passes", "max_depth", "jfa", "growth"):
            assert ra[k] == rb[k], (f, k, ra[k], rb[k])
        sa, sb = a.snapshot(), b.snapshot()
        for k in sa:
            assert np.array_equal(sa[k], sb[k]), (f, k)
        if f == 3:                        # a restore forces the frame-start rebuild
            a.restore(sa)
            b.restore(sb)
