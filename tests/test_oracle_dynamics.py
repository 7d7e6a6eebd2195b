"""Pins the oracle's hash grid and dynamics (grid.py, dynamics.py) on CPU."""
import math

import numpy as np
import pytest

import foam_inputs as fi

F = np.float32


# ---- grid.py: counting sort -------------------------------------------------

def cell_of(p, G, L=1.0):
    pw = p - F(L) * np.floor(p * F(1.0 / L) + F(0.5))
    q = ((pw + F(0.5 * L)) + F(1e-7)) * F(1.0 / L)
    c = np.mod(np.floor(q * F(G)).astype(np.int64), G)
    return (c[:, 0] * G + c[:, 1]) * G + c[:, 2]


def test_grid_is_stable_counting_sort(orc):
    n, G = 3000, 13
    pos, _ = fi.uniform(n, 0.01, 0.02, seed=5)
    pos[::7] += F(0.9)                       # some positions outside the box
    p, cc, cs, ci = orc.grid_build(pos, n, G, 1.0)
    cells = cell_of(pos, G)
    assert (cc == np.bincount(cells, minlength=G ** 3)).all()
    assert (cs == np.concatenate([[0], np.cumsum(cc)[:-1]])).all()
    assert (ci == np.argsort(cells, kind="stable")).all()
    # scatter stores wrapped positions back
    assert (p >= -0.5).all() and (p < 0.5).all()
    assert np.allclose(np.mod(p - pos + 0.5, 1.0), 0.5, atol=1e-6)


def test_rebuild_of_wrapped_positions_changes_nothing(orc):
    """grid.py:30-48 wrapP with L = 1 is idempotent in fp32, so a rebuild of
    positions a rebuild already wrapped reproduces positions and CSR bit for
    bit (the frame driver relies on it to reuse the end-of-frame grid).
    Exhaustive over every fp32 value with 0.25 <= |p| < 1 (where p + 0.5
    rounds), plus boundary values through the oracle's rebuild."""
    def wrap(p):
        return p - F(1.0) * np.floor(p * F(1.0) + F(0.5))
    lo, hi = F(0.25).view(np.uint32), F(1.0).view(np.uint32)
    u = np.arange(lo, hi, dtype=np.uint32)
    for sign in (0, 0x80000000):
        q = wrap((u | np.uint32(sign)).view(np.float32))
        assert (wrap(q).view(np.uint32) == q.view(np.uint32)).all()
    edge = np.array([0.5 - 2.0 ** -25, 0.5, -0.5, 0.5 - 2.0 ** -24, -0.5 - 2.0 ** -24, 1.5,
                     -1.5, 0.0, -0.0, 2.0 ** -30], np.float32)
    rng = np.random.default_rng(4)
    pos = np.concatenate([np.repeat(edge, 3),
                          ((rng.random(3 * 500) - 0.5) * 6).astype(F)]).reshape(-1, 3)
    n, G = pos.shape[0], 13
    p1, cc1, cs1, ci1 = orc.grid_build(pos, n, G, 1.0)
    p2, cc2, cs2, ci2 = orc.grid_build(p1, n, G, 1.0)
    assert (p1.view(np.uint32) == p2.view(np.uint32)).all()
    assert (cc1 == cc2).all() and (cs1 == cs2).all() and (ci1 == ci2).all()


# ---- dynamics.py:77-171 PBD ----------------------------------------------------

def pbd_args(cfg_gap=0.015):
    return float(F(1.0 + cfg_gap)), 0.2, 2.0


def test_pbd_two_particles_closed_form(orc):
    n, G = 2, 5
    r = F(0.05)
    d = F(0.08)                                  # overlap: ts = 0.1 * 1.015
    pos = np.array([[0.0, 0.0, 0.0], [d, 0.0, 0.0]], F)
    rad = np.array([r, r], F)
    p, cc, cs, ci = orc.grid_build(pos, n, G, 1.0)
    gap, mdf, dm = pbd_args()
    out = orc.pbd_jacobi(p, rad, n, G, 1.0, gap, mdf, dm, cs, cc, ci)
    ts = (r + r) * F(gap)
    dist = np.sqrt(d * d + F(1e-12))
    mag = min(F(0.5) * (ts - dist), F(0.2) * r)
    assert out[0, 0] == pytest.approx(-float(mag), rel=1e-6)
    assert out[1, 0] == pytest.approx(float(d + mag), rel=1e-6)
    assert (out[:, 1:] == 0).all()
    # Jacobi: symmetric pushes
    assert out[0, 0] + (out[1, 0] - d) == pytest.approx(0.0, abs=1e-8)


def test_pbd_push_is_capped_and_separated_pairs_untouched(orc):
    n, G = 3, 5
    r = F(0.05)
    pos = np.array([[0.0, 0.0, 0.0], [0.001, 0.0, 0.0], [0.3, 0.3, 0.3]], F)
    rad = np.array([r, r, r], F)
    p, cc, cs, ci = orc.grid_build(pos, n, G, 1.0)
    out = orc.pbd_jacobi(p, rad, n, G, 1.0, *pbd_args(), cs, cc, ci)
    assert out[0, 0] == pytest.approx(-0.2 * float(r), rel=1e-6)   # cap 0.2 r
    assert (out[2] == p[2]).all()


def test_pbd_across_periodic_boundary(orc):
    n, G = 2, 5
    r = F(0.05)
    pos = np.array([[-0.49, 0.0, 0.0], [0.48, 0.0, 0.0]], F)     # 0.03 apart via wrap
    rad = np.array([r, r], F)
    p, cc, cs, ci = orc.grid_build(pos, n, G, 1.0)
    out = orc.pbd_jacobi(p, rad, n, G, 1.0, *pbd_args(), cs, cc, ci)
    assert out[0, 0] > pos[0, 0] and out[1, 0] < pos[1, 0]


def test_gauss_seidel_differs_from_jacobi_only_in_order(orc):
    n, G = 500, 8
    pos, rad = fi.uniform(n, 0.02, 0.04, seed=9)
    p, cc, cs, ci = orc.grid_build(pos, n, G, 1.0)
    a = orc.pbd_jacobi(p, rad, n, G, 1.0, *pbd_args(), cs, cc, ci)
    b = orc.pbd_gauss_seidel(p, rad, n, G, 1.0, *pbd_args(), cs, cc, ci)
    assert (a[0] == b[0]).all()            # particle 0 sees the same snapshot
    assert not np.array_equal(a, b)


def test_max_overlap_pairs_counted_on_lower_index(orc):
    n, G = 3, 5
    pos = np.array([[0.0, 0.0, 0.0], [0.05, 0.0, 0.0], [0.3, 0.0, 0.0]], F)
    rad = np.array([0.04, 0.04, 0.01], F)
    p, cc, cs, ci = orc.grid_build(pos, n, G, 1.0)
    g, lmd = orc.max_overlap(p, rad, n, G, 1.0, cs, cc, ci)
    depth = F(0.08) - np.sqrt(F(0.05) * F(0.05) + F(1e-8))
    assert lmd[0] == pytest.approx(float(depth), rel=1e-6)
    assert lmd[1] == 0.0 and lmd[2] == 0.0      # only j > i contributes
    assert g == pytest.approx(float(depth), rel=1e-6)


# ---- dynamics.py:1242-1427 controller ------------------------------------------

def test_fsc_targets_branches(orc):
    fl, fh, gp, rmin, rmax = 10, 16, F(0.1), F(0.005), F(0.05)
    f_ema = np.array([0.0, 8.5, 9.5, 17.5, 16.5, 12.0, 14.0, 0.0], F)
    fsc = np.array([5, 0, 0, 0, 0, 0, 0, 20], np.int32)
    rad = np.full(8, 0.02, F)
    rt = orc.fsc_targets(rad, fsc, f_ema, 8, fl, fh, gp, rmin, rmax, 1.0)
    r0 = F(0.02)

    def ss(e0, e1, x):
        t = min(F(1), max(F(0), (F(x) - F(e0)) / (F(e1) - F(e0))))
        return F(t * t * (F(3) - F(2) * t))
    assert rt[0] == r0 * (F(1) + gp)                       # fsc fallback 5 < 9
    assert rt[1] == r0 * (F(1) + gp)                       # 8.5 < 9
    assert rt[2] == r0 * (F(1) + gp * (F(1) - ss(8, 10, 9.5)))
    assert rt[3] == r0 * (F(1) - gp)                       # 17.5 > 17
    assert rt[4] == r0 * (F(1) - gp * ss(16, 18, 16.5))
    assert rt[5] == r0 * (F(1) + gp * F(0.003))            # below centre 13
    assert rt[6] == r0 * (F(1) - gp * F(0.003))
    assert rt[7] == r0 * (F(1) - gp)                       # fsc 20


def test_nudge_radii_cap_backpressure_clamp(orc):
    rad = np.array([0.02, 0.02, 0.02, 0.06], F)
    rt = np.array([0.03, 0.01, 0.0201, 0.2], F)
    lmd = np.array([0.0, 0.0, 0.03, 0.0], F)
    a = F(1 - 0.05 ** (1 / 15))
    out = orc.nudge_radii(rad, rt, lmd, 4, a, F(0.12), F(0.005), F(0.05), 1.0, 2, F(0.035))
    cap = F(0.12) * F(0.02)
    assert out[0] == F(0.02) + min(a * (F(0.03) - F(0.02)), cap)
    assert out[1] == F(0.02) + max(a * (F(0.01) - F(0.02)), -cap)
    b = max(F(0.25), (F(0.035) - F(0.03)) / F(0.035))
    assert out[2] == F(0.02) + (a * (F(0.0201) - F(0.02))) * b
    assert out[3] == F(0.05)          # starts above r_max: final clamp binds


# ---- dynamics.py:1436-1557 pressure ----------------------------------------------

def test_lcg_hash_int32_wraparound(orc):
    for i, f in [(0, 0), (1, 0), (7, 3), (123456, 99), (2 ** 31 - 1, 5000)]:
        t = (i + 12345 * f) % 2 ** 32
        h = (1103515245 * t + 12345) % 2 ** 32
        assert orc.lcg_hash(i, f) == h & 0x7FFFFFFF


def test_pressure_two_particles_one_sided(orc):
    """face_ids row 0 lists 1 (lower row only); fsc counts both sides.  Only
    particle 0 changes: dV = clamp(k (V0 - V1), +-cap min V); r0' = (V0 -
    dV)^(1/3)."""
    n = 2
    rad = np.array([0.03, 0.02], F)
    ids = np.full((n, 32), -1, np.int32)
    ids[0, 0] = 1
    fsc = np.array([1, 1], np.int32)
    r, dr, mx, ch = orc.pressure(n, 0, rad, ids, fsc, F(0.10), F(3.0), F(0.02), 10,
                                 F(0.001), F(0.1))
    v0, v1 = np.float64(0.03) ** 3, np.float64(0.02) ** 3
    dv = max(-0.02 * min(v0, v1), min(0.02 * min(v0, v1), 0.10 * (v0 - v1)))
    assert r[0] == pytest.approx((v0 - dv) ** (1 / 3), rel=1e-5)
    assert r[1] == rad[1] and ch == 1
    assert mx == pytest.approx(abs(float(r[0]) - 0.03), rel=1e-3)


def test_pressure_budget_and_rotation(orc):
    """fsc=12 > max_nei=10: ten slots starting at hash % 12 (wrapping)."""
    n = 14
    rad = np.full(n, 0.02, F)
    rad[0] = 0.03
    ids = np.full((n, 32), -1, np.int32)
    ids[0, :12] = np.arange(1, 13)
    fsc = np.zeros(n, np.int32)
    fsc[0] = 12
    frame = 3
    start = orc.lcg_hash(0, frame) % 12
    used = {1 + (start + k) % 12 for k in range(10)}
    r, dr, _, _ = orc.pressure(n, frame, rad, ids, fsc, F(0.10), F(3.0), F(0.02), 10,
                               F(0.001), F(0.1))
    v0, v1 = np.float64(0.03) ** 3, np.float64(0.02) ** 3
    dv = min(0.02 * v1, 0.10 * (v0 - v1))
    per = (v0 - dv) ** (1 / 3) - 0.03
    assert dr[0] == pytest.approx(10 * per, rel=1e-4)
    assert len(used) == 10


# ---- noise ------------------------------------------------------------------------

@pytest.mark.parametrize("ctr,key,out", [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
])
def test_philox_known_answers(orc, ctr, key, out):
    """Random123 known-answer vectors for Philox4x32-10."""
    assert tuple(int(x) for x in orc.philox(ctr, key)) == out


def test_brownian_uniform_bounds_and_wrap(orc):
    n = 20000
    pos = np.zeros((n, 3), F)
    pos[0] = (0.49999, -0.5, 0.0)
    s = F(2e-4)
    out = orc.brownian(pos, n, s, 1.0, 7, 3, 0)
    d = out[1:] - pos[1:]
    assert (np.abs(d) <= s).all() and abs(d.mean()) < 1e-5
    assert d.std() == pytest.approx(float(s) / math.sqrt(3), rel=0.02)
    assert (out >= -0.5).all() and (out < 0.5).all()
    g = orc.brownian(pos, n, F(1e-4), 1.0, 7, 3, 1)[1:] - pos[1:]
    assert g.std() == pytest.approx(1e-4, rel=0.02)
    # keyed by (seed, frame): a different frame gives different noise
    assert not np.array_equal(out, orc.brownian(pos, n, s, 1.0, 7, 4, 0))


def test_topology_candidates_cap_and_order(orc):
    n, G = 40, 3
    pos = np.zeros((n, 3), F)
    pos[:, 0] = np.linspace(-0.05, 0.05, n, dtype=F)
    rad = np.full(n, 0.01, F)
    p, cc, cs, ci = orc.grid_build(pos, n, G, 1.0)
    pairs, cnt, tr = orc.topo_candidates(p, rad, n, G, 1.0, F(2.2), 28, cs, cc, ci)
    d = np.abs(p[None, :, 0] - p[:, None, 0])
    th = F(2.2) * F(0.02)
    for i in range(n):
        want = [j for j in ci if j > i and d[i, j] * d[i, j] <= th * th]
        assert list(pairs[i, :cnt[i]]) == want[:28]
        assert tr[i] == (len(want) > 28)


# ---- pins for the controller stats, Gaussian noise, OU jitter and XSPH ------------

def test_controller_stats_known_answers(orc):
    """Band [12, 16]: 11 grows, 16 and 12 are inside, 17 shrinks; dr known."""
    ema = np.array([11.0, 12.0, 16.0, 17.0, 3.0], F)
    rad = np.array([0.010, 0.020, 0.030, 0.040, 0.050], F)
    rt = np.array([0.011, 0.020, 0.027, 0.040, 0.058], F)
    oi, of = orc.controller_stats(rad, rt, ema, 5, 12, 16)
    assert list(oi) == [2, 1, 2]
    dr = [F(0.011) - F(0.010), F(0.0), F(0.027) - F(0.030), F(0.0), F(0.058) - F(0.050)]
    assert of[0] == pytest.approx(sum(float(x) for x in dr) / 5, rel=1e-6)
    assert of[1] == pytest.approx(float(F(0.058) - F(0.050)), rel=1e-6)


def test_gaussian_brownian_box_muller_from_philox(orc):
    """Axis x of particle i uses Philox(i, frame, 0, 0) words 0,1, y words 2,3,
    z Philox(i, frame, 1, 0) words 0,1, through Box-Muller in double; and the
    sample moments are those of N(0, sigma^2)."""
    seed, frame, s = 11, 5, F(1.0)
    n, L = 4, 1000.0
    pos = np.zeros((n, 3), F)
    out = orc.brownian(pos, n, s, L, seed, frame, 1)
    for i in range(n):
        r0 = orc.philox((i, frame, 0, 0), (seed, 0))
        r1 = orc.philox((i, frame, 1, 0), (seed, 0))

        def bm(a, b):
            u1 = ((int(a) >> 8) + 1) * 2.0 ** -24
            u2 = (int(b) >> 8) * 2.0 ** -24
            return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)
        want = [bm(r0[0], r0[1]), bm(r0[2], r0[3]), bm(r1[0], r1[1])]
        assert np.allclose(out[i], want, rtol=1e-6, atol=1e-7)
    n = 40000
    g = orc.brownian(np.zeros((n, 3), F), n, s, L, seed, frame, 1).astype(np.float64)
    assert abs(g.mean()) < 4.0 / math.sqrt(3 * n)
    assert g.std() == pytest.approx(1.0, rel=0.02)
    kurt = ((g - g.mean()) ** 4).mean() / g.var() ** 2
    assert kurt == pytest.approx(3.0, abs=0.15)
    c = np.corrcoef(g.T)
    assert np.abs(c - np.eye(3)).max() < 0.03


def test_ou_jitter_closed_forms(orc):
    n, dt, tau, rms = 20000, F(0.016), F(1.0), F(0.10)
    rad = np.full(n, 0.02, F)
    mean_r = F(0.02)
    # decay only (sigma 0): v' = v + (-v/tau)*dt, pos' = wrap(pos + v'*dt)
    v0 = np.full((n, 3), 0.01, F)
    pos = np.zeros((n, 3), F)
    p, v = orc.ou_jitter(pos, v0, rad, n, mean_r, dt, F(0.0), tau, F(1.0), 1.0, 3, 2)
    want_v = v0 + ((-v0 / tau) * dt)
    assert np.array_equal(v, want_v)
    assert np.allclose(p, want_v * dt, rtol=1e-6)
    # noise only: uniform increments of variance (sigma*sqrt(dt))^2 / 12
    p, v = orc.ou_jitter(pos, np.zeros((n, 3), F), rad, n, mean_r, dt, rms, tau, F(1.0),
                         1.0, 3, 2)
    sdt = float(rms) * float(mean_r) * math.sqrt(2.0 / float(tau)) * math.sqrt(float(dt))
    assert v.var() == pytest.approx(sdt * sdt / 12.0, rel=0.03)
    assert np.abs(v).max() <= sdt / 2 * (1 + 1e-6)
    # drift cap (dynamics.py:640-644): above cap = cap_frac * 2r the velocity
    # is scaled by (cap / step) / dt, so the step becomes cap / dt (the
    # reference divides by dt once more than a step-length cap would)
    big = np.full((4, 3), 5.0, F)
    cap_frac = F(0.2) * F(0.015)
    p, v = orc.ou_jitter(np.zeros((4, 3), F), big, rad[:4], 4, mean_r, dt, F(0.0), tau,
                         cap_frac, 1.0, 3, 2)
    step = np.sqrt((v.astype(np.float64) ** 2).sum(1)) * float(dt)
    cap = float(cap_frac) * 2 * 0.02
    assert np.allclose(step, cap / float(dt), rtol=1e-5)
    assert np.allclose(v, v[0], rtol=0) and (v > 0).all()      # direction kept


def test_xsph_two_particles_and_isolated(orc):
    """eps-blend with the mean velocity of neighbours inside lim: a moving
    particle next to a resting one; a third particle out of range keeps its
    velocity; zero velocities stay zero."""
    n, G, L = 3, 5, 1.0
    pos = np.array([[0.0, 0.0, 0.0], [0.01, 0.0, 0.0], [0.3, 0.3, 0.3]], F)
    vel = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 2.0, 0.0]], F)
    p, cc, cs, ci = orc.grid_build(pos, n, G, L)
    eps, lim = F(0.03), F(0.05 * 0.05)
    v = orc.xsph(p, vel, n, G, L, lim, eps, cs, cc, ci)
    assert np.allclose(v[0], [1.0 - 0.03, 0, 0], rtol=1e-6)
    assert np.allclose(v[1], [0.03, 0, 0], rtol=1e-6)
    assert np.allclose(v[2], [0, 2.0, 0], rtol=1e-6)
    assert (orc.xsph(p, np.zeros_like(vel), n, G, L, lim, eps, cs, cc, ci) == 0).all()
