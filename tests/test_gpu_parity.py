"""GPU parity: every libfos kernel against the CPU restatement on the same
seeded inputs, through the C ABI.

Bar: integer outputs (labels, filtered labels, face rows and counts,
neighbour lists, FSC, hash CSR) bit-exact; fp32 outputs bit-exact where the
two sides evaluate the same correctly rounded operations in the same order
(positions, radii, scores), otherwise within 1e-5 relative as stated per test.
"""
import ctypes
import math

import numpy as np
import pytest
import torch

import foam_config as fc
import foam_inputs as fi

pytestmark = pytest.mark.gpu
F = np.float32

# Device copies made inline in a call's argument list must outlive the call:
# a dropped temporary returns its block to torch's caching allocator, which
# hands it to the next copy, so two arguments would alias.
_alive = []


@pytest.fixture(autouse=True)
def _keep_device_copies():
    yield
    torch.cuda.synchronize()
    _alive.clear()


def dev(a):
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    _alive.append(t)
    return t


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


class G:
    """Thin numpy-in / numpy-out wrapper over the C ABI."""

    def __init__(self, fl, n, res, cells=27):
        self.fl = fl
        self.ctx = fl.Context(max(1, n), max(1, res), max(27, cells))
        self.h = self.ctx.handle

    def call(self, name, *a):
        self.fl.call(name, *a)

    def ptr(self, t):
        return self.fl.ptr(t)


def run_jfa_gpu(fl, pos, rad, n, res, passes, early=False, th=0.006, minp=3):
    g = G(fl, n, res)
    tp, tr = dev(pos), dev(rad)
    V = res ** 3
    grid = torch.empty(V, dtype=torch.int32, device="cuda")
    temp = torch.empty(V, dtype=torch.int32, device="cuda")
    info = torch.zeros(66, dtype=torch.int32, device="cuda")
    voxel = fc.f32(1.0 / res)
    st = fl.stream()
    g.call("fos_jfa_init", g.h, g.ptr(tp), g.ptr(tr), n, res, 1.0, voxel, 1.0, g.ptr(grid),
           g.ptr(temp), st)
    g.call("fos_jfa_passes", g.h, res, 1.0, voxel, passes, int(early), th, minp, g.ptr(grid),
           g.ptr(temp), g.ptr(info), st)
    return host(grid), host(info)


# ---- JFA ---------------------------------------------------------------------

def test_jfa_c3_full_size(gpu, orc):
    """The largest workload's JFA (N=200,000 clustered, 384^3, 10 passes;
    ids above 65,535, so the passes keep 32-bit candidate lists): labels and
    per-pass change counts bit-exact."""
    cfg = fc.preset("c3")
    pos, rad = fi.for_config(cfg)
    n, res, P = cfg.n, cfg.res, cfg.passes()
    print(f"[jfa parity] c3 n={n} res={res} passes={P}", flush=True)
    lab, passes, early, hist = orc.jfa_run(pos, rad, n, res, 1.0, fc.f32(1.0 / res), P)
    glab, info = run_jfa_gpu(gpu, pos, rad, n, res, P)
    assert passes == P and not early
    assert np.array_equal(glab, lab)
    assert info[0] == passes and (info[2:2 + passes] == hist).all()


@pytest.mark.parametrize("case", ["c1_64", "c1_48", "dense_seeds", "single"])
def test_jfa_init_and_passes_bit_exact(gpu, orc, case):
    if case == "c1_64":
        n, res = 2000, 64
        pos, rad = fi.uniform(n, 0.02, 0.04, seed=0)
    elif case == "c1_48":                         # non power of two
        n, res = 2000, 48
        pos, rad = fi.uniform(n, 0.02, 0.04, seed=1)
    elif case == "dense_seeds":                   # many seed collisions and ties
        n, res = 3000, 16
        pos, rad = fi.uniform(n, 0.01, 0.02, seed=2)
        pos[1::2] = pos[0::2]
        rad[1::2] = rad[0::2]
    else:
        n, res = 1, 20
        pos, rad = fi.uniform(1, 0.02, 0.04, seed=3)
    voxel = fc.f32(1.0 / res)
    P = int(math.ceil(math.log2(res))) + 1
    print(f"[jfa parity] case={case} n={n} res={res} passes={P}", flush=True)
    g0, _ = orc.jfa_init(pos, rad, n, res, 1.0, voxel)
    # seeding alone
    gg = G(gpu, n, res)
    grid = torch.empty(res ** 3, dtype=torch.int32, device="cuda")
    temp = torch.empty(res ** 3, dtype=torch.int32, device="cuda")
    gg.call("fos_jfa_init", gg.h, gg.ptr(dev(pos)), gg.ptr(dev(rad)), n, res, 1.0, voxel, 1.0,
            gg.ptr(grid), gg.ptr(temp), gpu.stream())
    assert (host(grid) == g0).all()
    assert (host(temp) == -1).all()
    lab, passes, early, hist = orc.jfa_run(pos, rad, n, res, 1.0, voxel, P)
    glab, info = run_jfa_gpu(gpu, pos, rad, n, res, P)
    assert (glab == lab).all()
    assert info[0] == passes and info[1] == 0 and (info[2:2 + passes] == hist).all()


def test_jfa_single_steps_match(gpu, orc):
    n, res = 700, 40
    pos, rad = fi.uniform(n, 0.02, 0.05, seed=4)
    voxel = fc.f32(1.0 / res)
    g = G(gpu, n, res)
    cur, _ = orc.jfa_init(pos, rad, n, res, 1.0, voxel)
    tp, tr = dev(pos), dev(rad)
    g.call("fos_jfa_seed_table", g.h, g.ptr(tp), g.ptr(tr), n, 1.0, gpu.stream())
    for step in (20, 10, 5, 2, 1, 1, 3):
        out, ch = orc.jfa_step(pos, rad, res, 1.0, voxel, step, cur)
        tin = dev(cur)
        tout = torch.empty_like(tin)
        chg = torch.zeros(1, dtype=torch.int32, device="cuda")
        g.call("fos_jfa_step", g.h, res, 1.0, voxel, step, g.ptr(tin), g.ptr(tout), g.ptr(chg),
               gpu.stream())
        assert (host(tout) == out).all(), step
        assert int(host(chg)[0]) == ch
        cur = out


# fos_set_param "jfa_variant": bit 0 four list entries per iteration (two
# otherwise), bit 1 the winning power distance carried from pass to pass,
# bit 2 the residue-class x-plane order
PASS_VARIANTS = {"batch2": 0, "batch4": 1, "batch2_carry": 2, "batch4_carry": 3,
                 "batch2_class": 4, "batch4_class": 5, "batch2_carry_class": 6,
                 "batch4_carry_class": 7}


@pytest.mark.parametrize("variant", sorted(PASS_VARIANTS))
def test_jfa_pass_kernel_variants(gpu, orc, variant):
    """Every pass-kernel variant gives the oracle's labels and change counts,
    including the collision-heavy dense case, a res that is not a multiple of
    64, more step-1 passes than the schedule (settled passes in a row), early
    exit, and ids above 65,535 (32-bit candidate lists)."""
    import ctypes
    v0 = ctypes.c_int()
    gpu.call("fos_get_param", b"jfa_variant", ctypes.byref(v0))
    try:
        gpu.call("fos_set_param", b"jfa_variant", PASS_VARIANTS[variant])
        cases = ((1500, 72, 7, False, 0, False), (3000, 16, 2, True, 0, False),
                 (900, 66, 5, False, 3, False), (2500, 48, 9, False, 0, True),
                 (70000, 96, 4, False, 0, False))
        for (n, res, seed, dup, extra, early) in cases:
            pos, rad = fi.uniform(n, 0.005, 0.03, seed=seed)
            if dup:
                pos[1::2] = pos[0::2]
                rad[1::2] = rad[0::2]
            voxel = fc.f32(1.0 / res)
            P = int(math.ceil(math.log2(res))) + 1 + extra
            print(f"[jfa variant] {variant} n={n} res={res} passes={P} early={early}",
                  flush=True)
            lab, passes, ex, hist = orc.jfa_run(pos, rad, n, res, 1.0, voxel, P, early)
            glab, info = run_jfa_gpu(gpu, pos, rad, n, res, P, early)
            assert (glab == lab).all()
            assert info[0] == passes and info[1] == int(ex)
            assert (info[2:2 + passes] == hist).all()
    finally:
        gpu.call("fos_set_param", b"jfa_variant", v0.value)


@pytest.mark.parametrize("orbits", [1, 0])
def test_jfa_half_pass_by_orbits(gpu, orc, orbits):
    """The step = res/2 pass by orbits (k_jfa_half, default) and by the list
    kernel give the oracle's labels and change counts: dense seeds (several
    owners per orbit), coincident particles (equal-distance ties between
    different ids), res 66 (step 33), and the distance carry (C2-size
    context) as well as 32-bit lists (ids above 65,535)."""
    import ctypes
    v0 = ctypes.c_int()
    gpu.call("fos_get_param", b"jfa_half_orbits", ctypes.byref(v0))
    try:
        gpu.call("fos_set_param", b"jfa_half_orbits", orbits)
        for (n, res, seed, dup) in ((20000, 32, 3, False), (3000, 16, 2, True), (900, 66, 5, False),
                                    (10000, 128, 0, False), (70000, 96, 4, False)):
            pos, rad = fi.uniform(n, 0.005, 0.03, seed=seed)
            if dup:
                pos[1::2] = pos[0::2]
                rad[1::2] = rad[0::2]
            voxel = fc.f32(1.0 / res)
            P = int(math.ceil(math.log2(res))) + 1
            print(f"[jfa half] orbits={orbits} n={n} res={res}", flush=True)
            lab, passes, ex, hist = orc.jfa_run(pos, rad, n, res, 1.0, voxel, P)
            glab, info = run_jfa_gpu(gpu, pos, rad, n, res, P)
            assert (glab == lab).all()
            assert info[0] == passes and (info[2:2 + passes] == hist).all()
            # the first pass alone, through the single-step entry point
            out, ch = orc.jfa_step(pos, rad, res, 1.0, voxel, res // 2,
                                   orc.jfa_init(pos, rad, n, res, 1.0, voxel)[0])
            g = G(gpu, n, res)
            tp, tr = dev(pos), dev(rad)
            tin = torch.empty(res ** 3, dtype=torch.int32, device="cuda")
            tout = torch.empty_like(tin)
            g.call("fos_jfa_init", g.h, g.ptr(tp), g.ptr(tr), n, res, 1.0, voxel, 1.0,
                   g.ptr(tin), None, gpu.stream())
            chg = torch.zeros(1, dtype=torch.int32, device="cuda")
            g.call("fos_jfa_step", g.h, res, 1.0, voxel, res // 2, g.ptr(tin), g.ptr(tout),
                   g.ptr(chg), gpu.stream())
            assert (host(tout) == out).all() and int(host(chg)[0]) == ch
    finally:
        gpu.call("fos_set_param", b"jfa_half_orbits", v0.value)


@pytest.mark.parametrize("res,extra", [(66, 3), (130, 2), (64, 4)])
def test_jfa_repeated_final_passes(gpu, orc, res, extra):
    """Several step-1 passes in a row: from the third on, a pass recomputes
    only voxels next to the previous pass's changes.  res 66 and 130 leave a
    partial last 64-voxel segment per row (z wrap across segments)."""
    n = 900
    pos, rad = fi.uniform(n, 0.01, 0.03, seed=res)
    voxel = fc.f32(1.0 / res)
    P = int(math.ceil(math.log2(res))) + 1 + extra
    print(f"[jfa repeated] res={res} passes={P}", flush=True)
    lab, passes, _, hist = orc.jfa_run(pos, rad, n, res, 1.0, voxel, P)
    glab, info = run_jfa_gpu(gpu, pos, rad, n, res, P)
    assert (glab == lab).all()
    assert info[0] == passes and (info[2:2 + passes] == hist).all()


def test_jfa_early_exit_on_device(gpu, orc):
    n, res = 2000, 64
    pos, rad = fi.uniform(n, 0.02, 0.04, seed=0)
    voxel = fc.f32(1.0 / res)
    _, _, _, hist = orc.jfa_run(pos, rad, n, res, 1.0, voxel, 7)
    th = float(hist[4]) / res ** 3 + 1e-9          # fires at pass 4 (or earlier)
    lab, passes, early, h2 = orc.jfa_run(pos, rad, n, res, 1.0, voxel, 7, True, th, 3)
    glab, info = run_jfa_gpu(gpu, pos, rad, n, res, 7, True, th, 3)
    assert early and info[1] == 1 and info[0] == passes
    assert (glab == lab).all()


def test_jfa_early_exit_fires_on_last_pass(gpu, orc):
    """Reference defaults (threshold 0.006, min 3 passes): the extra final
    step-1 pass changes almost nothing, so the exit fires on the last pass;
    all passes ran but early_exit is reported."""
    n, res = 2000, 64
    pos, rad = fi.uniform(n, 0.02, 0.04, seed=0)
    voxel = fc.f32(1.0 / res)
    lab, passes, early, hist = orc.jfa_run(pos, rad, n, res, 1.0, voxel, 7, True, 0.006, 3)
    assert early and passes == 7
    glab, info = run_jfa_gpu(gpu, pos, rad, n, res, 7, True, 0.006, 3)
    assert info[0] == 7 and info[1] == 1 and (info[2:9] == hist).all()
    assert (glab == lab).all()


def test_jfa_general_domain_size(gpu, orc):
    """L != 1 takes the division form of the minimum image."""
    n, res, L = 300, 32, 0.189
    pos, rad = fi.uniform(n, 0.004, 0.008, seed=5, L=L)
    voxel = fc.f32(L / res)
    P = 6
    lab, passes, _, hist = orc.jfa_run(pos, rad, n, res, L, voxel, P)
    g = G(gpu, n, res)
    grid = torch.empty(res ** 3, dtype=torch.int32, device="cuda")
    temp = torch.empty(res ** 3, dtype=torch.int32, device="cuda")
    info = torch.zeros(66, dtype=torch.int32, device="cuda")
    g.call("fos_jfa_init", g.h, g.ptr(dev(pos)), g.ptr(dev(rad)), n, res, L, voxel, 1.0,
           g.ptr(grid), None, gpu.stream())
    g.call("fos_jfa_passes", g.h, res, L, voxel, P, 0, 0.006, 3, g.ptr(grid), g.ptr(temp),
           g.ptr(info), gpu.stream())
    assert (host(grid) == lab).all()


# ---- majority + faces -----------------------------------------------------------

def gpu_faces(gpu, lab, res, n):
    g = G(gpu, n, res)
    tl = dev(lab)
    filt = torch.empty_like(tl)
    ids = torch.empty(n, 32, dtype=torch.int32, device="cuda")
    cts = torch.empty(n, 32, dtype=torch.int32, device="cuda")
    g.call("fos_majority", res, g.ptr(tl), g.ptr(filt), gpu.stream())
    g.call("fos_face_accumulate", g.h, res, n, g.ptr(filt), g.ptr(ids), g.ptr(cts), gpu.stream())
    return host(filt), host(ids), host(cts)


# (x-planes, z-extent) of the emission tile; x-planes 0 is the wave kernel
EMIT_TX = [(0, 64), (1, 64), (2, 64), (4, 64), (2, 32), (4, 16), (8, 16), (16, 16), (16, 32)]


def with_emit(gpu, shape, fn):
    tx, zw = shape
    return with_param(gpu, "face_emit_tx", tx,
                      lambda: with_param(gpu, "face_emit_zw", zw, fn))


@pytest.mark.parametrize("tx", EMIT_TX)
def test_majority_and_faces_c1(gpu, orc, tx):
    n, res = 2000, 64
    pos, rad = fi.uniform(n, 0.02, 0.04, seed=0)
    lab, _, _, _ = orc.jfa_run(pos, rad, n, res, 1.0, fc.f32(1 / 64), 7)
    f = orc.majority(lab, res)
    ids, cts = orc.accumulate_faces(f, res, n)
    gf, gids, gcts = with_emit(gpu, tx, lambda: gpu_faces(gpu, lab, res, n))
    assert (gf == f).all()
    assert (gids == ids).all() and (gcts == cts).all()


@pytest.mark.parametrize("tx", EMIT_TX)
def test_faces_row_truncation_and_table_growth(gpu, orc, tx):
    """Row 0 meets 100 distinct higher partners (> 32 kept, > 64 table
    slots): the first 32 by scan order survive with full counts."""
    res, n = 112, 101
    lab = np.zeros((res, res, res), np.int32)
    for k in range(100):
        lab[4 + k, :, 7] = 1 + k
    lab = lab.ravel()
    ids, cts = orc.accumulate_faces(lab, res, n)
    g = G(gpu, n, res)
    tl = dev(lab)
    gids = torch.empty(n, 32, dtype=torch.int32, device="cuda")
    gcts = torch.empty(n, 32, dtype=torch.int32, device="cuda")

    def run():
        g.call("fos_face_accumulate", g.h, res, n, g.ptr(tl), g.ptr(gids), g.ptr(gcts),
               gpu.stream())
        return host(gids), host(gcts)
    a, b = with_emit(gpu, tx, run)
    assert (a == ids).all() and (b == cts).all()
    # the grown tables keep working on the next call
    a, b = with_emit(gpu, tx, run)
    assert (a == ids).all() and (b == cts).all()


@pytest.mark.parametrize("tx", EMIT_TX)
def test_faces_random_labels(gpu, orc, tx):
    """Every voxel labelled at random: up to six distinct pairs per voxel
    overflow the workgroup merge table (pairs then go straight to the row
    tables) and each row meets hundreds of partners (the row tables grow)."""
    n, res = 1000, 32
    lab = np.random.default_rng(11).integers(-1, n, res ** 3).astype(np.int32)
    ids, cts = orc.accumulate_faces(lab, res, n)
    g = G(gpu, n, res)
    tl = dev(lab)
    gids = torch.empty(n, 32, dtype=torch.int32, device="cuda")
    gcts = torch.empty(n, 32, dtype=torch.int32, device="cuda")

    def run():
        g.call("fos_face_accumulate", g.h, res, n, g.ptr(tl), g.ptr(gids), g.ptr(gcts),
               gpu.stream())
        return host(gids), host(gcts)
    a, b = with_emit(gpu, tx, run)
    assert (a == ids).all() and (b == cts).all()


def test_faces_exclude_ids_at_or_above_n(gpu, orc):
    n, res = 500, 32
    pos, rad = fi.uniform(n, 0.02, 0.05, seed=7)
    lab, _, _, _ = orc.jfa_run(pos, rad, n, res, 1.0, fc.f32(1 / 32), 6)
    m = 400                                   # active_n below the labelled ids
    f = orc.majority(lab, res)
    ids, cts = orc.accumulate_faces(f, res, m)
    gf, gids, gcts = gpu_faces(gpu, lab, res, m)
    assert (gids == ids).all() and (gcts == cts).all()


# ---- finalize -----------------------------------------------------------------------

def gpu_finalize(gpu, n, rad, voxel, ids, cts, prev):
    g = G(gpu, n, 1)
    tp = dev(prev.astype(F))
    nl = torch.empty(n, 32, dtype=torch.int32, device="cuda")
    nc = torch.empty(n, dtype=torch.int32, device="cuda")
    used = torch.empty(n, dtype=torch.int32, device="cuda")
    fsc = torch.empty(n, dtype=torch.int32, device="cuda")
    g.call("fos_finalize_faces", g.h, n, g.ptr(dev(rad)), voxel, g.ptr(dev(ids)),
           g.ptr(dev(cts)), g.ptr(tp), g.ptr(nl), g.ptr(nc), g.ptr(used), g.ptr(fsc),
           gpu.stream())
    return host(tp), host(nl), host(nc), host(used), host(fsc)


def test_finalize_multi_frame_hysteresis(gpu, orc):
    """Several frames of C1-style faces with the hysteresis state carried
    over; includes many scores in the [0.10, 0.15) band."""
    n, res = 2000, 64
    voxel = fc.f32(1 / 64)
    pos, rad = fi.uniform(n, 0.02, 0.04, seed=0)
    rng = np.random.default_rng(3)
    prev_o = np.zeros((n, 32), F)
    prev_g = prev_o.copy()
    for frame in range(4):
        p = fi.wrap(pos + rng.normal(0, 0.004, pos.shape).astype(F))
        lab, _, _, _ = orc.jfa_run(p, rad, n, res, 1.0, voxel, 7)
        f = orc.majority(lab, res)
        ids, cts = orc.accumulate_faces(f, res, n)
        # push some counts into the hysteresis band
        band = rng.random(cts.shape) < 0.15
        cts = np.where(band & (cts > 0), np.maximum(1, (cts * 0.2).astype(np.int32)), cts)
        o = orc.finalize_faces(n, rad, voxel, ids, cts, prev_o)
        gout = gpu_finalize(gpu, n, rad, voxel, ids, cts, prev_g)
        for a, b, name in zip(o, gout, ("prev", "nl", "nc", "used", "fsc")):
            assert np.array_equal(a, b), (frame, name)
        prev_o, prev_g = o[0], gout[0]


def test_finalize_chains_and_overflow(gpu, orc):
    """Adversarial rows: ambiguous scores whose decisions depend on appends
    from lower rows (long dependency chains) and rows that overflow 32."""
    n = 300
    voxel = fc.f32(1 / 64)
    rad = np.full(n, 0.02, F)
    rng = np.random.default_rng(9)
    ids = np.full((n, 32), -1, np.int32)
    cts = np.zeros((n, 32), np.int32)
    for i in range(n - 1):
        k = rng.integers(1, 32)
        partners = np.sort(rng.choice(np.arange(i + 1, n), size=min(k, n - 1 - i), replace=False))
        rng.shuffle(partners)
        ids[i, :len(partners)] = partners
        cts[i, :len(partners)] = rng.choice([1, 1, 1, 2, 3], size=len(partners))
    for i in range(40):                       # row 250 overflows from lower rows
        row = [250] + [p for p in ids[i] if p not in (-1, 250)]
        ids[i, :] = -1
        ids[i, :min(32, len(row))] = row[:32]
        cts[i, 0] = 3
        cts[i, len(row):] = 0
    # rows list distinct partners, as accumulate_face_area produces them
    for i in range(n):
        v = ids[i][ids[i] >= 0]
        assert len(set(v.tolist())) == len(v)
    prev = (rng.random((n, 32)) < 0.3).astype(F) * F(0.12)
    o = orc.finalize_faces(n, rad, voxel, ids, cts, prev)
    gout = gpu_finalize(gpu, n, rad, voxel, ids, cts, prev)
    for a, b, name in zip(o, gout, ("prev", "nl", "nc", "used", "fsc")):
        assert np.array_equal(a, b), name


def test_finalize_long_chain(gpu, orc):
    """Row i's only slot (partner i + 1) has an ambiguous score, kept only
    when row i - 1 appended to it: the acceptance travels one row per round,
    so the fixed point takes ~n/2 round pairs -- several host bursts and more
    pairs than the change-flag ring holds.  Two calls on one context (the
    second sizes its first burst from the first)."""
    n = 150
    voxel = fc.f32(1 / 64)
    rad = np.full(n, 0.02, F)              # expected area 1.5 * (0.04 * 64)^2 = 9.83
    ids = np.full((n, 32), -1, np.int32)
    cts = np.zeros((n, 32), np.int32)
    ids[:n - 1, 0] = np.arange(1, n)
    cts[:n - 1, 0] = 1                     # score 0.102: ambiguous
    cts[0, 0] = 2                          # score 0.203: accepted
    prev = np.zeros((n, 32), F)
    o = orc.finalize_faces(n, rad, voxel, ids, cts, prev)
    # the chain reached the last row
    assert o[4][0] == 1 and (o[4][1:n - 1] == 2).all() and o[4][n - 1] == 1
    g = G(gpu, n, 1)
    for _ in range(2):
        tp = dev(prev)
        out = [torch.empty(n, 32, dtype=torch.int32, device="cuda")] + \
            [torch.empty(n, dtype=torch.int32, device="cuda") for _ in range(3)]
        g.call("fos_finalize_faces", g.h, n, g.ptr(dev(rad)), voxel, g.ptr(dev(ids)),
               g.ptr(dev(cts)), g.ptr(tp), *[g.ptr(t) for t in out], gpu.stream())
        gout = [host(tp)] + [host(t) for t in out]
        for a, b, name in zip(o, gout, ("prev", "nl", "nc", "used", "fsc")):
            assert np.array_equal(a, b), name


def test_ema_stats_validate(gpu, orc):
    n = 2000
    rng = np.random.default_rng(1)
    fsc = rng.integers(0, 20, n).astype(np.int32)
    ema = np.where(rng.random(n) < 0.3, 0.0, rng.random(n) * 15).astype(F)
    e = orc.fsc_ema(n, fc.f32(0.33), fsc, ema)
    te = dev(ema)
    gpu.call("fos_fsc_ema", n, 0.33, gpu.ptr(dev(fsc)), gpu.ptr(te), gpu.stream())
    assert np.array_equal(host(te), e)
    # stats / validate on a real C1 JFA frame
    res = 64
    pos, rad = fi.uniform(n, 0.02, 0.04, seed=0)
    voxel = fc.f32(1 / 64)
    lab, _, _, _ = orc.jfa_run(pos, rad, n, res, 1.0, voxel, 7)
    ids, cts = orc.accumulate_faces(orc.majority(lab, res), res, n)
    prev, nl, nc, used, fs = orc.finalize_faces(n, rad, voxel, ids, cts, np.zeros((n, 32), F))
    oi, of = orc.jfa_stats(n, rad, voxel, fs, e, used, ids, cts)
    g = G(gpu, n, res)
    goi = (gpu.I64 * 5)()
    gof = (gpu.D * 4)()
    g.call("fos_jfa_stats", g.h, n, g.ptr(dev(rad)), voxel, g.ptr(dev(fs)), g.ptr(dev(e)),
           g.ptr(dev(used)), g.ptr(dev(ids)), g.ptr(dev(cts)), goi, gof, gpu.stream())
    assert list(goi) == list(oi)
    assert np.allclose(list(gof), of, rtol=1e-9)
    vo = (gpu.I64 * 3)()
    g.call("fos_validate", g.h, n, g.ptr(dev(nl)), g.ptr(dev(nc)), vo, gpu.stream())
    assert list(vo) == list(orc.validate(n, nl, nc))


def test_dirty_tiles(gpu, orc):
    n, res = 2000, 64
    pos, rad = fi.uniform(n, 0.02, 0.04, seed=0)
    rng = np.random.default_rng(2)
    prev = fi.wrap(pos + rng.normal(0, 0.01, pos.shape).astype(F))
    rprev = rad * F(1.01)
    voxel = fc.f32(1 / 64)
    d = orc.mark_dirty_tiles(pos, rad, prev, rprev, n, res, 1.0, voxel)
    td = torch.zeros(64, dtype=torch.uint8, device="cuda")
    gpu.call("fos_dirty_mark", gpu.ptr(dev(pos)), gpu.ptr(dev(rad)), gpu.ptr(dev(prev)),
             gpu.ptr(dev(rprev)), n, res, 1.0, voxel, 16, 1, 0.5, 0.25, gpu.ptr(td), gpu.stream())
    assert np.array_equal(host(td), d)


# ---- grid + dynamics ---------------------------------------------------------------

def gpu_grid(gpu, pos, n, G_):
    g = G(gpu, n, 1, G_ ** 3)
    tp = dev(pos)
    cc = torch.empty(G_ ** 3, dtype=torch.int32, device="cuda")
    cs = torch.empty_like(cc)
    ci = torch.empty(n, dtype=torch.int32, device="cuda")
    g.call("fos_grid_build", g.h, g.ptr(tp), n, G_, 1.0, g.ptr(cc), g.ptr(cs), g.ptr(ci),
           gpu.stream())
    return tp, cc, cs, ci


def with_param(gpu, name, value, fn):
    import ctypes
    old = ctypes.c_int()
    gpu.call("fos_get_param", name.encode(), ctypes.byref(old))
    gpu.call("fos_set_param", name.encode(), value)
    try:
        return fn()
    finally:
        gpu.call("fos_set_param", name.encode(), old.value)


@pytest.mark.parametrize("block", [1, 0])
@pytest.mark.parametrize("n,G_,seed", [(2000, 13, 0), (10000, 21, 1), (5000, 3, 2), (1, 3, 6),
                                       (16256, 25, 7), (16257, 25, 8), (50000, 45, 9)])
def test_grid_build_bit_exact(gpu, orc, n, G_, seed, block):
    """Cell-range workgroups (block=1) and the multi-kernel path (block=0)
    give the oracle's grid, wrapped positions and in-cell ascending order,
    from one particle to 50,000 and from 27 cells to 91,125."""
    pos, _ = fi.uniform(n, 0.01, 0.02, seed=seed)
    pos[::5] += F(0.75)                       # unwrapped inputs
    print(f"[grid] n={n} G={G_} block={block}", flush=True)
    p, cc, cs, ci = orc.grid_build(pos, n, G_, 1.0)
    tp, tcc, tcs, tci = with_param(gpu, "grid_block", block, lambda: gpu_grid(gpu, pos, n, G_))
    assert np.array_equal(host(tp), p)
    assert np.array_equal(host(tcc), cc) and np.array_equal(host(tcs), cs)
    assert np.array_equal(host(tci), ci)


def test_grid_clustered_cells(gpu, orc):
    n, G_ = 20000, 34
    pos, _ = fi.clustered(n, seed=4)
    p, cc, cs, ci = orc.grid_build(pos, n, G_, 1.0)
    tp, tcc, tcs, tci = gpu_grid(gpu, pos, n, G_)
    assert np.array_equal(host(tci), ci) and np.array_equal(host(tcs), cs)


@pytest.mark.parametrize("block", [1, 0])
def test_grid_crowded_range(gpu, orc, block):
    """6000 particles in a few cells: a cell range holds more particles than
    its workgroup stages in LDS, and cells hold hundreds."""
    n, G_ = 6000, 34
    rng = np.random.default_rng(12)
    pos = (rng.random((n, 3)) * np.array([0.016, 0.06, 0.06]) + 0.1).astype(F)
    p, cc, cs, ci = orc.grid_build(pos, n, G_, 1.0)
    assert cc.max() > 300
    tp, tcc, tcs, tci = with_param(gpu, "grid_block", block, lambda: gpu_grid(gpu, pos, n, G_))
    assert np.array_equal(host(tp), p)
    assert np.array_equal(host(tcc), cc) and np.array_equal(host(tcs), cs)
    assert np.array_equal(host(tci), ci)


def test_pbd_maxoverlap_bit_exact(gpu, orc):
    n, G_ = 2000, 13
    pos, rad = fi.uniform(n, 0.02, 0.04, seed=0)
    p, cc, cs, ci = orc.grid_build(pos, n, G_, 1.0)
    gap = fc.f32(1.015)
    want = orc.pbd_jacobi(p, rad, n, G_, 1.0, gap, 0.2, 2.0, cs, cc, ci)
    tp, tcc, tcs, tci = gpu_grid(gpu, pos, n, G_)
    out = torch.empty_like(tp)
    tr = dev(rad)
    gpu.call("fos_pbd_pass", gpu.ptr(tp), gpu.ptr(out), gpu.ptr(tr), n, G_, 1.0, gap, 0.2, 2.0,
             gpu.ptr(tcs), gpu.ptr(tcc), gpu.ptr(tci), gpu.stream())
    got = host(out)
    assert np.allclose(got, want, rtol=1e-5, atol=1e-7)
    assert np.array_equal(got, want)
    g, lmd = orc.max_overlap(p, rad, n, G_, 1.0, cs, cc, ci)
    tl = torch.empty(n, dtype=torch.float32, device="cuda")
    tm = torch.empty(1, dtype=torch.float32, device="cuda")
    gpu.call("fos_max_overlap", gpu.ptr(tp), gpu.ptr(tr), n, G_, 1.0, gpu.ptr(tcs), gpu.ptr(tcc),
             gpu.ptr(tci), gpu.ptr(tl), gpu.ptr(tm), gpu.stream())
    assert np.array_equal(host(tl), lmd) and float(host(tm)[0]) == g


@pytest.mark.parametrize("pr", [False, 8, 16, 32])
@pytest.mark.parametrize("block", [1, 0])
def test_pbd_pass_cells_equals_pass_then_rebuild(gpu, orc, block, pr):
    """The PBD pass that also runs the next rebuild's wrap-and-hash step,
    finished by fos_grid_build_from_cells, equals fos_pbd_pass followed by
    a full rebuild (the oracle's), including particles pushed across the
    box faces (wrapped twice, as the reference does).  pr (lanes per
    particle): the pass reads slot records written by fos_grid_build_pr, and
    the finishing rebuild writes the next ones (checked against the CSR)."""
    n, G_ = 3000, 14
    pos, rad = fi.uniform(n, 0.02, 0.04, seed=5)
    pos[:40, 0] = F(0.4999)                   # crowd the +x face
    pos[40:80, 1] = F(-0.4999)
    p, cc, cs, ci = orc.grid_build(pos, n, G_, 1.0)
    gap = fc.f32(1.015)
    moved = orc.pbd_jacobi(p, rad, n, G_, 1.0, gap, 0.2, 2.0, cs, cc, ci)
    want = orc.grid_build(moved, n, G_, 1.0)

    def run():
        g = G(gpu, n, 1, G_ ** 3)
        tp = dev(pos)
        tr = dev(rad)
        tcc = torch.empty(G_ ** 3, dtype=torch.int32, device="cuda")
        tcs = torch.empty_like(tcc)
        tci = torch.empty(n, dtype=torch.int32, device="cuda")
        rec = torch.full((n, 4), -7.0, dtype=torch.float32, device="cuda") if pr else None
        g.call("fos_grid_build_pr", g.h, g.ptr(tp), g.ptr(tr), n, G_, 1.0, g.ptr(tcc),
               g.ptr(tcs), g.ptr(tci), g.ptr(rec) if pr else None, gpu.stream())
        out = torch.empty_like(tp)
        g.call("fos_pbd_pass_cells", g.h, g.ptr(tp), g.ptr(out), g.ptr(tr), n, G_, 1.0,
               gap, 0.2, 2.0, g.ptr(tcs), g.ptr(tcc), g.ptr(tci), g.ptr(rec) if pr else None,
               gpu.stream())
        g.call("fos_grid_build_from_cells", g.h, g.ptr(out), g.ptr(tr), n, G_, g.ptr(tcc),
               g.ptr(tcs), g.ptr(tci), g.ptr(rec) if pr else None, gpu.stream())
        return host(out), host(tcc), host(tcs), host(tci), None if rec is None else host(rec)
    *got, rec = with_param(gpu, "grid_block", block,
                           lambda: with_param(gpu, "pbd_lanes", pr or 0, run))
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    if pr:
        ci = want[3]
        assert np.array_equal(rec[:, :3], want[0][ci]) and np.array_equal(rec[:, 3], rad[ci])


@pytest.mark.parametrize("lanes", [8, 16, 32])
def test_pbd_crowded_hits_take_several_windows(gpu, orc, lanes):
    """A crowded cluster: particles with more corrections than one LDS window
    (48) add them in several windows; the pass (slot records, any lanes per
    particle) equals the oracle's serial Jacobi pass bit for bit."""
    n, G_ = 3000, 10
    rng = np.random.default_rng(11)
    pos, rad = fi.uniform(n, 0.02, 0.04, seed=7)
    pos[:600] = (rng.normal(0.0, 0.03, (600, 3))).astype(F)      # a dense ball
    rad[:600] = F(0.045)
    p, cc, cs, ci = orc.grid_build(pos, n, G_, 1.0)
    gap = fc.f32(1.015)
    moved = orc.pbd_jacobi(p, rad, n, G_, 1.0, gap, 0.2, 2.0, cs, cc, ci)
    want = orc.grid_build(moved, n, G_, 1.0)
    # the cluster really overflows one window (count the hits on the CPU)
    d = p[:600, None, :] - p[None, :600, :]
    d -= np.round(d)
    dist = np.sqrt((d * d).sum(-1) + F(1e-12))
    hits = ((dist < (rad[:600, None] + rad[None, :600]) * gap).sum(1) - 1).max()
    assert hits > 2 * 48, hits

    def run():
        g = G(gpu, n, 1, G_ ** 3)
        tp, tr = dev(pos), dev(rad)
        tcc = torch.empty(G_ ** 3, dtype=torch.int32, device="cuda")
        tcs = torch.empty_like(tcc)
        tci = torch.empty(n, dtype=torch.int32, device="cuda")
        rec = torch.empty((n, 4), dtype=torch.float32, device="cuda")
        g.call("fos_grid_build_pr", g.h, g.ptr(tp), g.ptr(tr), n, G_, 1.0, g.ptr(tcc),
               g.ptr(tcs), g.ptr(tci), g.ptr(rec), gpu.stream())
        out = torch.empty_like(tp)
        g.call("fos_pbd_pass_cells", g.h, g.ptr(tp), g.ptr(out), g.ptr(tr), n, G_, 1.0,
               gap, 0.2, 2.0, g.ptr(tcs), g.ptr(tcc), g.ptr(tci), g.ptr(rec), gpu.stream())
        g.call("fos_grid_build_from_cells", g.h, g.ptr(out), g.ptr(tr), n, G_, g.ptr(tcc),
               g.ptr(tcs), g.ptr(tci), None, gpu.stream())
        return host(out), host(tcc), host(tcs), host(tci)
    got = with_param(gpu, "pbd_lanes", lanes, run)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("case", ["c1", "overflow"])
def test_jfa_faces_equals_accumulate_then_finalize(gpu, orc, case):
    """fos_jfa_faces (emission, compaction and finalize behind one host read)
    against the oracle's accumulate_face_area + finalize_faces, twice on one
    context.  'overflow': row 0 meets 100 partners, so the first call's row
    tables overflow, grow and the emission starts over."""
    if case == "c1":
        n, res = 2000, 64
        pos, rad = fi.uniform(n, 0.02, 0.04, seed=0)
        lab, _, _, _ = orc.jfa_run(pos, rad, n, res, 1.0, fc.f32(1 / 64), 7)
        lab = orc.majority(lab, res)
    else:
        res, n = 112, 101
        lab = np.zeros((res, res, res), np.int32)
        for k in range(100):
            lab[4 + k, :, 7] = 1 + k
        lab = lab.ravel()
        rad = np.full(n, 0.02, F)
    voxel = fc.f32(1.0 / res)
    rng = np.random.default_rng(8)
    prev = ((rng.random((n, 32)) < 0.3) * F(0.12)).astype(F)
    ids, cts = orc.accumulate_faces(lab, res, n)
    want = orc.finalize_faces(n, rad, voxel, ids, cts, prev)
    g = G(gpu, n, res)
    for call in range(2):
        tp = dev(prev)
        fid = torch.empty(n, 32, dtype=torch.int32, device="cuda")
        fct = torch.empty(n, 32, dtype=torch.int32, device="cuda")
        out = [torch.empty(n, 32, dtype=torch.int32, device="cuda")] + \
            [torch.empty(n, dtype=torch.int32, device="cuda") for _ in range(3)]
        g.call("fos_jfa_faces", g.h, res, n, g.ptr(dev(lab)), g.ptr(dev(rad)), voxel, g.ptr(fid),
               g.ptr(fct), g.ptr(tp), *[g.ptr(t) for t in out], gpu.stream())
        assert np.array_equal(host(fid), ids) and np.array_equal(host(fct), cts), call
        got = [host(tp)] + [host(t) for t in out]
        for a, b, name in zip(want, got, ("prev", "nl", "nc", "used", "fsc")):
            assert np.array_equal(a, b), (call, name)


def test_fsc_gate_and_gated_targets(gpu, orc):
    """fos_fsc_gate: 0 while frame < warm start or more than zero_rate of
    the particles have FSC 0, else 1 (run.py:1184-1195), on consecutive
    calls (its counters reset themselves); fos_fsc_targets_gated leaves the
    targets alone on 0 and equals fos_fsc_targets on 1."""
    g = G(gpu, 5000, 1)
    gate = torch.full((1,), -1.0, dtype=torch.float32, device="cuda")
    rng = np.random.default_rng(4)
    for n, zeros, frame, warm, rate in ((5000, 100, 30, 20, 0.05), (5000, 400, 30, 20, 0.05),
                                        (3000, 0, 5, 20, 0.05), (5000, 250, 30, 20, 0.05),
                                        (1, 1, 30, 20, 0.05), (0, 0, 30, 20, 0.05)):
        fsc = rng.integers(1, 25, max(n, 1)).astype(np.int32)[:n]
        fsc[rng.permutation(n)[:zeros]] = 0
        f0 = zeros / max(1, n)
        want = 0.0 if (frame < warm or f0 > rate) else 1.0
        g.call("fos_fsc_gate", g.h, g.ptr(dev(fsc)) if n else None, n, frame, warm, rate,
               g.ptr(gate), gpu.stream())
        assert float(host(gate)[0]) == want, (n, zeros, frame)
    n = 4000
    rad = rng.uniform(0.01, 0.04, n).astype(F)
    fsc = rng.integers(0, 25, n).astype(np.int32)
    ema = (rng.random(n) * 25).astype(F)
    rt = orc.fsc_targets(rad, fsc, ema, n, 12, 16, fc.f32(0.1), fc.f32(0.006), fc.f32(0.024))
    for gv in (0.0, 1.0):
        trt = dev(np.full(n, 7.0, F))
        gt = dev(np.array([gv], F))
        g.call("fos_fsc_targets_gated", g.ptr(dev(rad)), g.ptr(trt), g.ptr(dev(fsc)),
               g.ptr(dev(ema)), n, 12, 16, 0.1, 0.006, 0.024, 1.0, g.ptr(gt), gpu.stream())
        assert np.array_equal(host(trt), rt if gv else np.full(n, 7.0, F)), gv


def test_controller_kernels(gpu, orc):
    n = 5000
    rng = np.random.default_rng(6)
    rad = rng.uniform(0.01, 0.04, n).astype(F)
    fsc = rng.integers(0, 25, n).astype(np.int32)
    ema = np.where(rng.random(n) < 0.2, 0.0, rng.random(n) * 25).astype(F)
    rt = orc.fsc_targets(rad, fsc, ema, n, 12, 16, fc.f32(0.1), fc.f32(0.006), fc.f32(0.024))
    trt = torch.empty(n, dtype=torch.float32, device="cuda")
    gpu.call("fos_fsc_targets", gpu.ptr(dev(rad)), gpu.ptr(trt), gpu.ptr(dev(fsc)),
             gpu.ptr(dev(ema)), n, 12, 16, 0.1, 0.006, 0.024, 1.0, gpu.stream())
    assert np.array_equal(host(trt), rt)
    lmd = (rng.random(n) * 0.05).astype(F)
    a = fc.f32(1 - 0.05 ** (1 / 15))
    for mode in (0, 1, 2):
        want = orc.nudge_radii(rad, rt, lmd, n, a, fc.f32(0.12), fc.f32(0.006), fc.f32(0.024),
                               fc.f32(0.6), mode, fc.f32(0.035))
        tr = dev(rad)
        gpu.call("fos_nudge_radii", gpu.ptr(tr), gpu.ptr(trt), gpu.ptr(dev(lmd)), n, a, 0.12,
                 0.006, 0.024, 0.6, mode, 0.035, gpu.stream())
        assert np.array_equal(host(tr), want), mode
    oi, of = orc.controller_stats(rad, rt, ema, n, 12, 16)
    g = G(gpu, n, 1)
    goi = (gpu.I64 * 3)()
    gof = (gpu.D * 2)()
    g.call("fos_controller_stats", g.h, gpu.ptr(dev(rad)), gpu.ptr(trt), gpu.ptr(dev(ema)), n, 12,
           16, goi, gof, gpu.stream())
    assert list(goi) == list(oi) and gof[1] == of[1]
    assert gof[0] == pytest.approx(of[0], rel=1e-9, abs=1e-15)


def test_pressure_iterations(gpu, orc):
    n, res = 2000, 64
    voxel = fc.f32(1 / 64)
    pos, rad = fi.uniform(n, 0.02, 0.04, seed=0)
    lab, _, _, _ = orc.jfa_run(pos, rad, n, res, 1.0, voxel, 7)
    ids, cts = orc.accumulate_faces(orc.majority(lab, res), res, n)
    _, _, _, _, fs = orc.finalize_faces(n, rad, voxel, ids, cts, np.zeros((n, 32), F))
    r = rad.copy()
    tr = dev(rad)
    tdr = torch.empty(n, dtype=torch.float32, device="cuda")
    st = torch.zeros(2, dtype=torch.float32, device="cuda")
    tids, tfs = dev(ids), dev(fs)
    for frame in range(10):
        r, dr, mx, ch = orc.pressure(n, frame, r, ids, fs, fc.f32(0.1), fc.f32(3.0),
                                     fc.f32(0.02), 10, fc.f32(0.01), fc.f32(0.04))
        gpu.call("fos_pressure_iter", n, frame, gpu.ptr(tr), gpu.ptr(tdr), gpu.ptr(tids),
                 gpu.ptr(tfs), 0.1, 3.0, 0.02, 10, 0.01, 0.04, gpu.ptr(st), gpu.stream())
        got = host(tr)
        assert np.allclose(got, r, rtol=1e-5, atol=0)
        s = host(st)
        assert s.view(np.int32)[1] == ch
        assert s[0] == pytest.approx(mx, rel=1e-5)
    # the double-evaluated power makes the radii bit-identical in practice
    assert np.array_equal(host(tr), r)


@pytest.mark.parametrize("n,iters,max_nei", [(2000, 20, 10), (2000, 7, 24), (50000, 3, 10)])
def test_pressure_run(gpu, orc, n, iters, max_nei):
    """fos_pressure_run (all iterations of a frame) equals the oracle's
    iterations with the same frame; the stats are the last iteration's."""
    res = 64 if n <= 2000 else 128
    voxel = fc.f32(1 / res)
    rmin, rmax = (0.02, 0.04) if n <= 2000 else (0.008, 0.016)
    pos, rad = fi.uniform(n, rmin, rmax, seed=3)
    lab, _, _, _ = orc.jfa_run(pos, rad, n, res, 1.0, voxel, 8)
    ids, cts = orc.accumulate_faces(orc.majority(lab, res), res, n)
    _, _, _, _, fs = orc.finalize_faces(n, rad, voxel, ids, cts, np.zeros((n, 32), F))
    frame = 5
    r = rad.copy()
    for _ in range(iters):
        r, dr, mx, ch = orc.pressure(n, frame, r, ids, fs, fc.f32(0.1), fc.f32(3.0),
                                     fc.f32(0.02), max_nei, fc.f32(rmin), fc.f32(rmax))
    ctx = gpu.Context(n, 1, 27)
    tr = dev(rad)
    tdr = torch.empty(n, dtype=torch.float32, device="cuda")
    st = torch.zeros(2, dtype=torch.float32, device="cuda")
    print(f"[pressure_run] n={n} iters={iters}", flush=True)
    gpu.call("fos_pressure_run", ctx.handle, n, frame, iters, gpu.ptr(tr), gpu.ptr(tdr),
             gpu.ptr(dev(ids)), gpu.ptr(dev(fs)), 0.1, 3.0, 0.02, max_nei, rmin, rmax,
             gpu.ptr(st), gpu.stream())
    assert np.array_equal(host(tr), r)
    assert np.array_equal(host(tdr), dr)
    s = host(st)
    assert s.view(np.int32)[1] == ch and s[0] == pytest.approx(mx, rel=1e-6)
    del ctx


def test_pressure_one_launch_equals_oracle(gpu, orc):
    """The one-launch pressure iterations (per-partner waits, no barrier)
    give the oracle's radii bit for bit, call after call on one context:
    a repeated call, a call with other radii and frame, a smaller n with a
    ragged last block, two iterations, and the per-launch path beside it
    (pressure_flow 0)."""
    n, res = 3000, 64
    voxel = fc.f32(1 / res)
    pos, rad = fi.uniform(n, 0.02, 0.04, seed=11)
    lab, _, _, _ = orc.jfa_run(pos, rad, n, res, 1.0, voxel, 8)
    ids, cts = orc.accumulate_faces(orc.majority(lab, res), res, n)
    _, _, _, _, fs = orc.finalize_faces(n, rad, voxel, ids, cts, np.zeros((n, 32), F))
    ctx = gpu.Context(n, 1, 27)
    v = ctypes.c_int()
    gpu.call("fos_get_param", b"pressure_flow", ctypes.byref(v))
    assert v.value == 1
    rng = np.random.default_rng(4)
    other = (rad * rng.uniform(0.9, 1.1, n)).astype(F)
    # (n, frame, iterations, radii, budget): budgets 16, 10, 7, 1 and 0 (a
    # group's lanes past the budget add +0)
    cases = [(n, 5, 20, rad, 10), (n, 5, 20, rad, 10), (n, 9, 20, other, 10),
             (1237, 2, 13, rad, 10), (n, 3, 2, rad, 10), (n, 4, 32, rad, 10),
             (n, 6, 6, other, 16), (1237, 8, 5, rad, 7), (n, 7, 3, rad, 1), (n, 2, 3, rad, 0)]
    try:
        for flow in (1, 0):
            gpu.call("fos_set_param", b"pressure_flow", flow)
            for m, frame, iters, r0, budget in cases:
                # rows of the first m particles that point past m are skipped
                # the same way by the oracle and the kernels (j >= n)
                r = np.clip(r0[:m], fc.f32(0.02), fc.f32(0.04)).astype(F)
                want = r.copy()
                for _ in range(iters):
                    want, dr, mx, ch = orc.pressure(m, frame, want, ids[:m], fs[:m],
                                                    fc.f32(0.1), fc.f32(3.0), fc.f32(0.02), budget,
                                                    fc.f32(0.02), fc.f32(0.04))
                tr = dev(r)
                tdr = torch.empty(m, dtype=torch.float32, device="cuda")
                st = torch.zeros(2, dtype=torch.float32, device="cuda")
                print(f"[pressure_flow] flow={flow} n={m} frame={frame} iters={iters} "
                      f"budget={budget}", flush=True)
                gpu.call("fos_pressure_run", ctx.handle, m, frame, iters, gpu.ptr(tr),
                         gpu.ptr(tdr), gpu.ptr(dev(ids[:m])), gpu.ptr(dev(fs[:m])), 0.1, 3.0,
                         0.02, budget, 0.02, 0.04, gpu.ptr(st), gpu.stream())
                assert np.array_equal(host(tr), want), (flow, m, frame, iters, budget)
                assert np.array_equal(host(tdr), dr), (flow, m, frame, iters, budget)
                s = host(st)
                assert s.view(np.int32)[1] == ch and s[0] == pytest.approx(mx, rel=1e-6)
    finally:
        gpu.call("fos_set_param", b"pressure_flow", 1)
    with pytest.raises(gpu.FosError):
        gpu.call("fos_set_param", b"pressure_flow", 2)
    del ctx


def test_pressure_one_launch_large_foam(gpu, orc):
    """A foam above what the 16-lane one-launch form holds at once (N =
    40,000: one lane per particle) against the oracle and the per-launch
    path, on synthetic face rows with skipped entries (-1, self) and
    budgets from 0 to the 32-entry row."""
    n = 40000
    rng = np.random.default_rng(21)
    fs = rng.integers(0, 33, n).astype(np.int32)
    fs[::97] = 0
    ids = np.full((n, 32), -1, np.int32)
    for i in range(n):
        c = int(fs[i])
        ids[i, :c] = rng.choice(n, c, replace=False)
    ids[::13, 0] = np.arange(0, n, 13)      # a self entry
    ids[5::17, 1] = -1                      # a hole inside a row
    rad = rng.uniform(0.02, 0.04, n).astype(F)
    want = rad.copy()
    for _ in range(20):
        want, dr, mx, ch = orc.pressure(n, 3, want, ids, fs, fc.f32(0.1), fc.f32(3.0),
                                        fc.f32(0.02), 10, fc.f32(0.02), fc.f32(0.04))
    ctx = gpu.Context(n, 1, 27)
    try:
        for flow in (1, 0):
            gpu.call("fos_set_param", b"pressure_flow", flow)
            tr = dev(rad)
            tdr = torch.empty(n, dtype=torch.float32, device="cuda")
            st = torch.zeros(2, dtype=torch.float32, device="cuda")
            print(f"[pressure_flow large] flow={flow}", flush=True)
            gpu.call("fos_pressure_run", ctx.handle, n, 3, 20, gpu.ptr(tr), gpu.ptr(tdr),
                     gpu.ptr(dev(ids)), gpu.ptr(dev(fs)), 0.1, 3.0, 0.02, 10, 0.02, 0.04,
                     gpu.ptr(st), gpu.stream())
            assert np.array_equal(host(tr), want), flow
            assert np.array_equal(host(tdr), dr), flow
            s = host(st)
            assert s.view(np.int32)[1] == ch and s[0] == pytest.approx(mx, rel=1e-6)
    finally:
        gpu.call("fos_set_param", b"pressure_flow", 1)
    del ctx


def test_noise_kernels(gpu, orc):
    n = 10000
    pos, rad = fi.uniform(n, 0.012, 0.024, seed=0)
    for mode, s in ((0, 2e-4), (1, 1e-4)):
        want = orc.brownian(pos, n, fc.f32(s), 1.0, 12345, 17, mode)
        tp = dev(pos)
        gpu.call("fos_brownian", gpu.ptr(tp), n, s, 1.0, 12345, 17, mode, gpu.stream())
        got = host(tp)
        assert np.allclose(got, want, rtol=1e-5, atol=1e-9)
        assert (got == want).mean() > 0.999
    v = np.random.default_rng(1).normal(0, 1e-3, (n, 3)).astype(F)
    wp, wv = orc.ou_jitter(pos, v, rad, n, fc.f32(0.018), fc.f32(0.016), fc.f32(0.1),
                           fc.f32(1.0), fc.f32(0.003), 1.0, 99, 4)
    tp, tv = dev(pos), dev(v)
    gpu.call("fos_ou_jitter", gpu.ptr(tp), gpu.ptr(tv), gpu.ptr(dev(rad)), n, 0.018, 0.016, 0.1,
             1.0, 0.003, 1.0, 99, 4, gpu.stream())
    assert np.array_equal(host(tv), wv) and np.array_equal(host(tp), wp)
    # the mean radius computed on the device (run.py:860-861): the exact
    # double sum rounded to fp32, so the OU step matches the oracle's frame
    # (which takes f32(mean(rad, float64))) bit for bit, for ragged n too
    g = G(gpu, n, 1)
    for m in (n, 1, 1023, 4097):
        r = rad[:m].copy()
        mean = F(np.mean(r, dtype=np.float64))
        wp, wv = orc.ou_jitter(pos[:m], v[:m], r, m, mean, fc.f32(0.016), fc.f32(0.1),
                               fc.f32(1.0), fc.f32(0.003), 1.0, 99, 4)
        tp, tv = dev(pos[:m]), dev(v[:m])
        tm = torch.zeros(1, dtype=torch.float32, device="cuda")
        gpu.call("fos_ou_jitter_mean", g.h, gpu.ptr(tp), gpu.ptr(tv), gpu.ptr(dev(r)), m, 0.016,
                 0.1, 1.0, 0.003, 1.0, 99, 4, gpu.ptr(tm), gpu.stream())
        assert host(tm)[0] == mean, m
        assert np.array_equal(host(tv), wv) and np.array_equal(host(tp), wp), m


def test_brownian_cells_equals_brownian_then_rebuild(gpu):
    """fos_brownian_cells + fos_grid_build_from_cells (with slot records)
    equals fos_brownian + fos_grid_build_pr bit for bit, both noise modes,
    with particles pushed across the box faces."""
    n, G_ = 5000, 17
    pos, rad = fi.uniform(n, 0.01, 0.03, seed=21)
    pos[:50, 0] = F(0.49995)
    pos[50:100, 2] = F(-0.5)
    for mode, s in ((0, 2e-4), (1, 1e-4)):
        outs = []
        for fused in (False, True):
            g = G(gpu, n, 1, G_ ** 3)
            tp, tr = dev(pos), dev(rad)
            cc = torch.empty(G_ ** 3, dtype=torch.int32, device="cuda")
            cs, ci = torch.empty_like(cc), torch.empty(n, dtype=torch.int32, device="cuda")
            rec = torch.empty(n, 4, dtype=torch.float32, device="cuda")
            st = gpu.stream()
            if fused:
                g.call("fos_brownian_cells", g.h, g.ptr(tp), n, s, 1.0, 777, 5, mode, G_, st)
                g.call("fos_grid_build_from_cells", g.h, g.ptr(tp), g.ptr(tr), n, G_, g.ptr(cc),
                       g.ptr(cs), g.ptr(ci), g.ptr(rec), st)
            else:
                g.call("fos_brownian", g.ptr(tp), n, s, 1.0, 777, 5, mode, st)
                g.call("fos_grid_build_pr", g.h, g.ptr(tp), g.ptr(tr), n, G_, 1.0, g.ptr(cc),
                       g.ptr(cs), g.ptr(ci), g.ptr(rec), st)
            outs.append([host(t) for t in (tp, cc, cs, ci, rec)])
        for a, b in zip(*outs):
            assert np.array_equal(a.view(np.uint32) if a.dtype == F else a,
                                  b.view(np.uint32) if b.dtype == F else b), mode


def test_jfa_passes_counters_after_init_and_alone(gpu, orc):
    """fos_jfa_init clears the pass counters for the fos_jfa_passes call that
    follows; a second fos_jfa_passes call without an init clears them itself.
    Both give the oracle's labels and per-pass change counts (early exit on,
    so stale counts would also change the schedule)."""
    n, res = 2000, 48
    pos, rad = fi.uniform(n, 0.02, 0.04, seed=1)
    voxel = fc.f32(1.0 / res)
    P = int(math.ceil(math.log2(res))) + 1
    lab, passes, early, hist = orc.jfa_run(pos, rad, n, res, 1.0, voxel, P, True, 0.006, 3)
    g = G(gpu, n, res)
    st = gpu.stream()
    grid = torch.empty(res ** 3, dtype=torch.int32, device="cuda")
    temp = torch.empty_like(grid)
    g.call("fos_jfa_init", g.h, g.ptr(dev(pos)), g.ptr(dev(rad)), n, res, 1.0, voxel, 1.0,
           g.ptr(grid), g.ptr(temp), st)
    seeded = grid.clone()
    for call in range(3):
        grid.copy_(seeded)
        temp.fill_(-1)
        info = torch.zeros(66, dtype=torch.int32, device="cuda")
        g.call("fos_jfa_passes", g.h, res, 1.0, voxel, P, 1, 0.006, 3, g.ptr(grid), g.ptr(temp),
               g.ptr(info), st)
        inf = host(info)
        assert np.array_equal(host(grid), lab), call
        assert inf[0] == passes and inf[1] == int(early), call
        assert (inf[2:2 + passes] == hist).all(), call


def test_settle_nudge_counter_gate_and_backpressure(gpu, orc):
    """fos_settle_nudge against the host rule of run.py:733-751 and 1231
    with the oracle's nudge: the counter entering the frame (forced, stored,
    or reset by the last frame's gate), the countdown it writes, and the
    global backpressure (mode 1) computed from the device overlap maximum in
    double as the host computes it."""
    n, adj = 3000, 15
    rng = np.random.default_rng(12)
    rad = rng.uniform(0.006, 0.024, n).astype(F)
    rt = rng.uniform(0.006, 0.024, n).astype(F)
    lmd = (rng.random(n) * 0.05).astype(F)
    a = fc.f32(1 - 0.05 ** (1 / 15))
    tol = fc.f32(0.035)
    for mode in (1, 2):
        for s_in, force, apply_gate, gate, md in ((3, 0, 0, 0.0, 0.01), (0, 0, 1, 1.0, 0.03),
                                                  (0, 0, 1, 0.0, 0.0), (0, 1, 0, 0.0, 0.05),
                                                  (1, 0, 1, 0.0, 0.002), (5, 0, 1, 1.0, 0.02)):
            s = adj if force else s_in
            if apply_gate and gate > 0.5:
                s = adj
            back = max(0.25, min(1.0, (float(tol) - float(F(md))) / float(tol)))
            want = (orc.nudge_radii(rad, rt, lmd, n, a, fc.f32(0.12), fc.f32(0.006),
                                    fc.f32(0.024), fc.f32(back), mode, tol) if s > 0 else rad)
            tr = dev(rad)
            settle = torch.tensor([s_in, -99], dtype=torch.int32, device="cuda")
            maxd = torch.tensor([md, gate], dtype=torch.float32, device="cuda")
            gpu.call("fos_settle_nudge", gpu.ptr(tr), gpu.ptr(dev(rt)), gpu.ptr(dev(lmd)), n, a,
                     0.12, 0.006, 0.024, mode, tol, float(tol), gpu.ptr(maxd), gpu.ptr(settle),
                     gpu.ptr(settle[1:]), adj, force, apply_gate, gpu.stream())
            case = (mode, s_in, force, apply_gate, gate)
            assert np.array_equal(host(tr), want), case
            assert int(host(settle)[1]) == (s - 1 if s > 0 else s), case


def test_topology_candidates(gpu, orc):
    n, G_ = 3000, 10
    pos, rad = fi.uniform(n, 0.01, 0.03, seed=8)
    p, cc, cs, ci = orc.grid_build(pos, n, G_, 1.0)
    pairs, cnt, tr = orc.topo_candidates(p, rad, n, G_, 1.0, fc.f32(2.2), 28, cs, cc, ci)
    tp, tcc, tcs, tci = gpu_grid(gpu, pos, n, G_)
    gp = torch.full((n, 28), -1, dtype=torch.int32, device="cuda")
    gc = torch.empty(n, dtype=torch.int32, device="cuda")
    gt = torch.empty(n, dtype=torch.int32, device="cuda")
    gpu.call("fos_topo_candidates", gpu.ptr(tp), gpu.ptr(dev(rad)), n, G_, 1.0, 2.2, 28,
             gpu.ptr(tcs), gpu.ptr(tcc), gpu.ptr(tci), gpu.ptr(gp), gpu.ptr(gc), gpu.ptr(gt),
             gpu.stream())
    assert np.array_equal(host(gc), cnt) and np.array_equal(host(gt), tr)
    assert np.array_equal(host(gp), pairs)


def test_topology_degree_kernels(gpu, orc):
    """Gabriel (full and batched), k-NN, degree EMA and radius rule against
    the oracle (topology.py:140-504)."""
    n, G, cap = 3000, 12, 28
    pos, rad = fi.uniform(n, 0.005, 0.02, seed=12)
    p, cc, cs, ci = orc.grid_build(pos, n, G, 1.0)
    pairs, cnt, tr = orc.topo_candidates(p, rad, n, G, 1.0, F(2.2), cap, cs, cc, ci)
    ema = np.random.default_rng(3).uniform(10, 25, n).astype(F)
    tp, tcs, tcc, tci = dev(p), dev(cs), dev(cc), dev(ci)
    deg = torch.empty(n, dtype=torch.int32, device="cuda")
    for stride, early in ((1, 0), (3, 1)):
        want = orc.topo_gabriel(p, n, G, 1.0, 1.0 / G, cs, cc, ci, pairs, cnt, cap, 7, stride,
                                bool(early), 18.0, ema)
        gpu.call("fos_topo_gabriel", gpu.ptr(tp), n, G, 1.0, 1.0 / G, gpu.ptr(tcs), gpu.ptr(tcc),
                 gpu.ptr(tci), gpu.ptr(dev(pairs)), gpu.ptr(dev(cnt)), cap, 7, stride, early, 18.0,
                 gpu.ptr(dev(ema)), gpu.ptr(deg), gpu.stream())
        assert np.array_equal(host(deg), want), (stride, early)
    B, q = 4, 3
    wdeg = np.zeros(n, np.int32)
    gdeg = torch.zeros(n, dtype=torch.int32, device="cuda")
    gpairs = torch.full((n, cap), -1, dtype=torch.int32, device="cuda")
    gcnt = torch.zeros(n, dtype=torch.int32, device="cuda")
    gtr = torch.zeros(n, dtype=torch.int32, device="cuda")
    trd = dev(rad)
    for b in range(B):
        wp, wc, wt, wdeg = orc.topo_gabriel_batched(p, rad, n, G, 1.0, 1.0 / G, F(2.2), cap, cs,
                                                    cc, ci, b, B, q, 11, wdeg)
        gpu.call("fos_topo_gabriel_batched", gpu.ptr(tp), gpu.ptr(trd), n, G, 1.0, 1.0 / G, 2.2,
                 cap, gpu.ptr(tcs), gpu.ptr(tcc), gpu.ptr(tci), b, B, q, 11, gpu.ptr(gpairs),
                 gpu.ptr(gcnt), gpu.ptr(gtr), gpu.ptr(gdeg), gpu.stream())
        mine = np.arange(n) % B == b
        gc = host(gcnt)
        assert np.array_equal(gc[mine], wc[mine]) and np.array_equal(host(gtr)[mine], wt[mine])
        gpp = host(gpairs)
        for i in np.nonzero(mine)[0]:
            m = min(wc[i], cap)
            assert (gpp[i, :m] == wp[i, :m]).all()
    assert np.array_equal(host(gdeg), wdeg)
    kd = orc.topo_knn(p, rad, n, G, 1.0, F(2.2), 14, cs, cc, ci)
    gpu.call("fos_topo_knn", gpu.ptr(tp), gpu.ptr(trd), n, G, 1.0, 2.2, 14, gpu.ptr(tcs),
             gpu.ptr(tcc), gpu.ptr(tci), gpu.ptr(deg), gpu.stream())
    assert np.array_equal(host(deg), kd)
    te = dev(ema)
    gpu.call("fos_topo_ema", gpu.ptr(dev(kd)), gpu.ptr(te), n, 0.1, gpu.stream())
    assert np.array_equal(host(te), orc.topo_ema(kd, ema, n, F(0.1)))
    de = np.random.default_rng(4).uniform(5, 25, n).astype(F)
    tr_ = dev(rad)
    gpu.call("fos_topo_radii", gpu.ptr(tr_), gpu.ptr(dev(de)), n, 10.0, 18.0, 0.05, 0.05, 0.005,
             0.02, gpu.stream())
    assert np.array_equal(host(tr_), orc.topo_radii(rad, de, n, 10.0, 18.0, F(0.05), F(0.05),
                                                     F(0.005), F(0.02)))
