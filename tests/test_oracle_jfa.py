"""Pins the oracle's JFA path (jfa.py) without the reference runtime.

Each check comes from the mathematics or from a hand-derived case of the
reference's serial loop, chosen so a plausible slip (wrong sign or offset
order, non-strict comparison, swapped index, missing wrap) fails it.
"""
import numpy as np
import pytest

import foam_inputs as fi

F = np.float32


def centres(res, voxel):
    c = (np.arange(res, dtype=F) + F(0.5)) * F(voxel) - F(0.5)
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(F)


def exact_power_labels(pos, rad, res):
    """Brute-force argmin of |c-p|^2 - r^2 over all particles (min image)."""
    C = centres(res, 1.0 / res)
    best = np.full(res ** 3, np.inf, F)
    arg = np.full(res ** 3, -1, np.int64)
    for i in range(len(rad)):
        d = C - pos[i]
        d = d - np.round(d)
        pd = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2] - F(rad[i]) * F(rad[i])
        m = pd < best
        best[m] = pd[m]
        arg[m] = i
    return arg


# ---- jfa.py:451-477 power distance ----------------------------------------

def test_power_distance_closed_form(orc):
    # voxel centre at origin, particle at (0.1, -0.2, 0.3), r = 0.05
    pd = orc.power_distance((0.0, 0.0, 0.0), (0.1, -0.2, 0.3), 0.05)
    d = np.array([-0.1, 0.2, -0.3], F)
    expect = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2] - F(0.05) * F(0.05)
    assert pd == pytest.approx(float(expect), rel=1e-6)
    # beta scales the radius term
    pd2 = orc.power_distance((0.0, 0.0, 0.0), (0.1, -0.2, 0.3), 0.05, beta=2.0)
    assert pd2 == pytest.approx(0.14 - 0.01, rel=1e-5)


def test_power_distance_minimum_image(orc):
    # 0.45 and -0.45 are 0.1 apart through the periodic boundary
    pd = orc.power_distance((0.45, 0.0, 0.0), (-0.45, 0.0, 0.0), 0.0)
    assert pd == pytest.approx(0.01, rel=1e-5)
    # exact half-box separation: both images have the same square
    pd_half = orc.power_distance((0.25, 0.0, 0.0), (-0.25, 0.0, 0.0), 0.0)
    assert pd_half == F(0.25)


def test_seed_voxel_formula(orc):
    res = 64
    # floor(((p + 0.5) / 1) * 64), linear x*R^2 + y*R + z
    assert orc.seed_voxel((-0.5, -0.5, -0.5), res) == 0
    assert orc.seed_voxel((0.0, 0.0, 0.0), res) == (32 * res + 32) * res + 32
    x, y, z = 0.49999, -0.49, 0.1
    ix, iy, iz = [int(np.floor((F(v) + F(0.5)) * F(res))) for v in (x, y, z)]
    assert orc.seed_voxel((x, y, z), res) == (ix * res + iy) * res + iz
    # a coordinate that rounds up to +L/2 wraps to voxel 0
    assert orc.seed_voxel((0.5, 0.0, 0.0), res) == (0 * res + 32) * res + 32


# ---- jfa.py:485-520 seeding ------------------------------------------------

def test_seed_collision_lower_power_distance_wins(orc):
    res, voxel = 8, F(1 / 8)
    # both in voxel (4,4,4) whose centre is (0.0625, 0.0625, 0.0625)
    pos = np.array([[0.10, 0.10, 0.10], [0.07, 0.07, 0.07]], F)
    rad = np.array([0.01, 0.01], F)
    g, t = orc.jfa_init(pos, rad, 2, res, 1.0, voxel)
    v = (4 * res + 4) * res + 4
    assert g[v] == 1                      # particle 1 is closer to the centre
    assert (g >= 0).sum() == 1 and (t == -1).all()
    # a larger radius lowers the power distance: particle 0 now wins
    rad2 = np.array([0.09, 0.01], F)
    g2, _ = orc.jfa_init(pos, rad2, 2, res, 1.0, voxel)
    assert g2[v] == 0


def test_seed_collision_tie_keeps_lowest_index(orc):
    res, voxel = 8, F(1 / 8)
    pos = np.array([[0.3, 0.3, 0.3], [0.1, 0.1, 0.1], [0.1, 0.1, 0.1], [0.1, 0.1, 0.1]], F)
    rad = np.array([0.02, 0.02, 0.02, 0.02], F)
    g, _ = orc.jfa_init(pos, rad, 4, res, 1.0, voxel)
    v = orc.seed_voxel(pos[1], res)
    assert g[v] == 1


# ---- jfa.py:524-623 one pass ------------------------------------------------

def test_step_scan_order_breaks_ties(orc):
    """Two candidates at equal power distance: the one met first in the
    dx, dy, dz scan (dx = -1 before dx = +1) wins; the incumbent wins ties."""
    res, voxel, step = 16, F(1 / 16), 4
    c = (np.arange(res, dtype=F) + F(0.5)) * voxel - F(0.5)
    x0 = 8
    # particles centred on voxels x0-4 and x0+4 (same y, z), equal radii
    pos = np.array([[c[x0 + 4], c[8], c[8]], [c[x0 - 4], c[8], c[8]]], F)
    rad = np.array([0.01, 0.01], F)
    lab = np.full(res ** 3, -1, np.int32)
    lab[((x0 + 4) * res + 8) * res + 8] = 0
    lab[((x0 - 4) * res + 8) * res + 8] = 1
    out, ch = orc.jfa_step(pos, rad, res, 1.0, voxel, step, lab)
    v = (x0 * res + 8) * res + 8
    assert out[v] == 1        # dx = -1 (particle 1 at x0-4) is scanned first
    lab2 = lab.copy()
    lab2[v] = 0               # incumbent 0 at equal distance is kept
    out2, _ = orc.jfa_step(pos, rad, res, 1.0, voxel, step, lab2)
    assert out2[v] == 0
    # changes count the voxels whose label differs from the input
    assert ch == int((out != lab).sum())


def test_single_particle_floods_everything(orc):
    res = 16
    pos = np.array([[0.123, -0.321, 0.4]], F)
    rad = np.array([0.05], F)
    lab, passes, early, hist = orc.jfa_run(pos, rad, 1, res, 1.0, F(1 / res), 5)
    assert passes == 5 and not early
    assert (lab == 0).all()


@pytest.mark.parametrize("res,n", [(16, 12), (24, 30)])
def test_jfa_close_to_exact_power_diagram(orc, res, n):
    pos, rad = fi.uniform(n, 0.02, 0.08, seed=res)
    lab, passes, _, _ = orc.jfa_run(pos, rad, n, res, 1.0, F(1 / res),
                                    int(np.ceil(np.log2(res))) + 1)
    exact = exact_power_labels(pos, rad, res)
    assert (lab == exact).mean() > 0.98


def test_two_equal_particles_bisector(orc):
    """Equal radii: the boundary is the bisector plane x = 0 (and the
    periodic one at x = +-0.5); labels follow x exactly."""
    res = 16
    pos = np.array([[-0.25, 0.0, 0.0], [0.25, 0.0, 0.0]], F)
    rad = np.array([0.03, 0.03], F)
    lab, _, _, _ = orc.jfa_run(pos, rad, 2, res, 1.0, F(1 / res), 5)
    x = np.arange(res ** 3) // (res * res)
    assert (lab[x < 8] == 0).all() and (lab[x >= 8] == 1).all()


def test_early_exit_rule(orc):
    pos, rad = fi.uniform(40, 0.02, 0.06, seed=3)
    res = 32
    full, p_full, e_full, hist = orc.jfa_run(pos, rad, 40, res, 1.0, F(1 / res), 6)
    assert p_full == 6 and not e_full
    V = res ** 3
    # exit after the first pass q >= 2 whose change fraction is below th
    th = float(hist[3]) / V + 1e-12
    _, p, e, h2 = orc.jfa_run(pos, rad, 40, res, 1.0, F(1 / res), 6, early_exit=True, th=th,
                              min_passes=3)
    q = next(k for k in range(2, 6) if hist[k] / V < th)
    assert e and p == q + 1 and (h2 == hist[:p]).all()


# ---- jfa.py:640-699 majority filter ------------------------------------------

def test_majority_island_and_ties(orc):
    res = 4
    g = np.zeros(res ** 3, np.int32) + 5
    v = lambda x, y, z: (x * res + y) * res + z
    g[v(1, 1, 1)] = 9                       # island of one voxel -> 5
    g[v(3, 3, 3)] = -1                      # unassigned stays -1
    out = orc.majority(g, res)
    assert out[v(1, 1, 1)] == 5
    assert out[v(3, 3, 3)] == -1
    assert out[v(0, 0, 0)] == 5
    # tie: centre 7 (count 1), neighbours 2,2 (x-1,x+1), 3,3 (y-1,y+1), 4,4 (z)
    g = np.zeros(res ** 3, np.int32)
    g[v(1, 1, 1)] = 7
    g[v(0, 1, 1)] = g[v(2, 1, 1)] = 2
    g[v(1, 0, 1)] = g[v(1, 2, 1)] = 3
    g[v(1, 1, 0)] = g[v(1, 1, 2)] = 4
    out = orc.majority(g, res)
    assert out[v(1, 1, 1)] == 2             # first label reaching the max count
    # centre keeps its label against single distinct neighbours
    g = np.arange(res ** 3, dtype=np.int32)
    out = orc.majority(g, res)
    assert (out == g).all()


# ---- jfa.py:703-762 face accumulation ------------------------------------------

def faces_reference_sets(lab, res, n):
    """Set-based restatement: every 6-face (I, I+e) with 0 <= a=lab[I] < b=
    lab[I+e] < n counts once for pair (a, b); its scan key is
    6*lin(I) + 2*axis + (dir > 0); rows list pairs by their smallest key."""
    L3 = lab.reshape(res, res, res)
    pairs = {}
    for axis in range(3):
        for k, d in enumerate((-1, 1)):
            nb = np.roll(L3, -d, axis=axis).ravel()
            a = lab
            key = 6 * np.arange(res ** 3, dtype=np.int64) + 2 * axis + k
            m = (a >= 0) & (a < n) & (nb >= 0) & (nb < n) & (a < nb)
            for aa, bb, kk in zip(a[m], nb[m], key[m]):
                c, k0 = pairs.get((aa, bb), (0, kk))
                pairs[(aa, bb)] = (c + 1, min(k0, kk))
    rows = {}
    for (a, b), (c, k) in pairs.items():
        rows.setdefault(a, []).append((k, b, c))
    ids = np.full((n, 32), -1, np.int32)
    cts = np.zeros((n, 32), np.int32)
    for a, lst in rows.items():
        lst.sort()
        for s, (k, b, c) in enumerate(lst[:32]):
            ids[a, s], cts[a, s] = b, c
    return ids, cts


def test_face_accumulation_matches_set_definition(orc):
    res, n = 20, 25
    pos, rad = fi.uniform(n, 0.03, 0.09, seed=11)
    lab, _, _, _ = orc.jfa_run(pos, rad, n, res, 1.0, F(1 / res), 6)
    filt = orc.majority(lab, res)
    ids, cts = orc.accumulate_faces(filt, res, n)
    eids, ects = faces_reference_sets(filt, res, n)
    assert (ids == eids).all() and (cts == ects).all()


def test_face_accumulation_truncates_row_at_32(orc):
    """Row 0 with 40 higher partners keeps the first 32 met in scan order, with
    full counts; the rest vanish."""
    res, n = 48, 41
    lab = np.zeros((res, res, res), np.int32)
    # 40 slabs of other labels stacked along y at x = 2 .. 41, z = 0
    for k in range(40):
        lab[2 + k, :, 5] = 1 + k
    lab = lab.ravel()
    ids, cts = orc.accumulate_faces(lab, res, n)
    eids, ects = faces_reference_sets(lab, res, n)
    assert (ids == eids).all() and (cts == ects).all()
    assert (ids[0] >= 0).sum() == 32 and ids[0, 0] == 1 and ids[0, 31] == 32


def test_face_accumulation_ignores_ids_at_or_above_n(orc):
    res = 6
    lab = np.zeros(res ** 3, np.int32)
    lab[:res * res] = 3   # x = 0 plane is label 3
    ids, cts = orc.accumulate_faces(lab, res, 3)     # n = 3 excludes label 3
    assert (ids == -1).all() and (cts == 0).all()
    ids, cts = orc.accumulate_faces(lab, res, 4)
    # faces between the x=0 plane and the x=1 / x=5 planes: 2 * 36
    assert ids[0, 0] == 3 and cts[0, 0] == 2 * res * res


# ---- jfa.py:789-857 finalize -----------------------------------------------------

def expected_area(ri, rj, voxel):
    t = (F(2.0) * (F(0.5) * (F(ri) + F(rj)))) / F(voxel)
    return F(1.5) * (t * t)


def test_finalize_known_answers(orc):
    """Four particles, voxel 1/64, equal radii 0.02: expected area
    1.5 * (0.04*64)^2 = 9.8304; count 2 -> 0.203 (accept), 1 -> 0.1017
    (hysteresis band), 0 counts are skipped."""
    n, voxel = 4, F(1 / 64)
    rad = np.full(n, 0.02, F)
    ids = np.full((n, 32), -1, np.int32)
    cts = np.zeros((n, 32), np.int32)
    ids[0, :3] = [1, 2, 3]
    cts[0, :3] = [2, 1, 1]      # (0,1) accept; (0,2), (0,3) in the band
    ids[1, :2] = [2, 3]
    cts[1, :2] = [1, 0]         # (1,2) band; (1,3) has no voxels -> skipped
    prev = np.zeros((n, 32), F)
    prev[0, 2] = 0.12           # slot 2 of row 0 was kept last frame
    prev[1, 1] = 0.2            # row 1 position 1 (its slot for partner 3)
    e = expected_area(0.02, 0.02, voxel)
    s2, s1 = F(2) / e, F(1) / e
    assert s2 >= F(0.15) and F(0.10) <= s1 < F(0.15)
    p, nl, nc, used, fsc = orc.finalize_faces(n, rad, voxel, ids, cts, prev)
    # row 0: slot 0 accept -> append 1 at position 0 (prev[0,0] = s2);
    # slot 1 band, prev[0,1]=0 and position 1 not yet written -> reject,
    # prev[0,1] = 0; slot 2 band, prev[0,2] = 0.12 -> keep, append 3 at 1.
    assert list(nl[0, :2]) == [1, 3] and fsc[0] == 2
    assert p[0, 0] == s2 and p[0, 1] == s1 and p[0, 2] == F(0.12)
    # row 1 got 0 appended at position 0 before its turn; slot 0 (partner 2)
    # is in the band and reads prev[1,0] = s2 written by row 0 -> keep.
    assert list(nl[1, :2]) == [0, 2] and fsc[1] == 2 and p[1, 1] == s1
    # row 2 received 1 only; row 3 received 0
    assert list(nl[2, :1]) == [1] and fsc[2] == 1
    assert list(nl[3, :1]) == [0] and fsc[3] == 1
    assert (nc == used).all() and (used == fsc).all()


def test_finalize_row_overflow_is_asymmetric(orc):
    """Rows 1..32 each accept a pair with 33, filling row 33 before its own
    turn; 33's own accept of 34 then cannot be listed in row 33 (fsc stays
    32) while row 34 still lists 33."""
    n, voxel = 40, F(1 / 64)
    rad = np.full(n, 0.02, F)
    ids = np.full((n, 32), -1, np.int32)
    cts = np.zeros((n, 32), np.int32)
    ids[1:33, 0] = 33
    cts[1:33, 0] = 5
    ids[33, 0] = 34
    cts[33, 0] = 5
    prev = np.zeros((n, 32), F)
    p, nl, nc, used, fsc = orc.finalize_faces(n, rad, voxel, ids, cts, prev)
    assert fsc[33] == 32 and list(nl[33]) == list(range(1, 33))
    assert fsc[34] == 1 and nl[34, 0] == 33
    assert all(fsc[k] == 1 and nl[k, 0] == 33 for k in range(1, 33))
    v = orc.validate(n, nl, nc)
    assert list(v) == [1, 1, 0]          # 34 -> 33 unreturned, row 33 full


def test_ema_rule(orc):
    fsc = np.array([0, 4, 4, 7], np.int32)
    ema = np.array([0.0, 0.0, 2.0, 6.0], F)
    out = orc.fsc_ema(4, F(0.33), fsc, ema)
    a = F(0.33)
    assert out[0] == F(0.0) and out[1] == F(4.0)
    assert out[2] == a * F(4) + (F(1) - a) * F(2.0)
    assert out[3] == a * F(7) + (F(1) - a) * F(6.0)


def test_validate_counts_asymmetry_and_self_loops(orc):
    n = 3
    nl = np.full((n, 32), -1, np.int32)
    nc = np.zeros(n, np.int32)
    nl[0, 0] = 1          # 0 -> 1 but 1 does not list 0
    nl[2, 0] = 2          # self loop (also counts as symmetric: 2 lists 2)
    out = orc.validate(n, nl, nc)
    assert list(out) == [1, 0, 1]


def test_jfa_stats_quirk_scores_from_face_rows(orc):
    n, voxel = 2, F(1 / 64)
    rad = np.full(n, 0.02, F)
    ids = np.full((n, 32), -1, np.int32)
    cts = np.zeros((n, 32), np.int32)
    ids[0, 0], cts[0, 0] = 1, 3
    fsc = np.array([1, 1], np.int32)
    used = np.array([1, 1], np.int32)
    oi, of = orc.jfa_stats(n, rad, voxel, fsc, np.array([1.0, 1.0], F), used, ids, cts)
    assert list(oi) == [2, 0, 1, 1, 2]
    assert of[2] == pytest.approx(float(F(3) / expected_area(0.02, 0.02, voxel)))
    assert of[3] == 0.0


def test_dirty_tiles_marking(orc):
    res, L = 64, 1.0
    voxel = F(1 / 64)
    pos = np.array([[0.0, 0.0, 0.0], [0.2, 0.2, 0.2]], F)
    prev = pos.copy()
    prev[1, 0] += F(0.02)                 # moved > 0.5 voxel
    rad = np.array([0.02, 0.02], F)
    d = orc.mark_dirty_tiles(pos, rad, prev, rad, 2, res, L, voxel)
    tpa = 4
    assert d.sum() == 27                  # one tile + halo 1 (all distinct)
    t = [(int((F(0.2) + F(0.5)) * F(64))) // 16] * 3
    assert d[(t[0] * tpa + t[1]) * tpa + t[2]] == 1
