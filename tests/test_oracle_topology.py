"""Pins the oracle's topology kernels (topology.py:140-504) on CPU: the
Gabriel witness test against brute force, the hash subsampling, the batched
form against the full one, the k-NN proxy, the degree EMA and the radius
rule."""
import numpy as np
import pytest

import foam_inputs as fi

F = np.float32


def gw(p):
    return (p - np.floor(p + F(0.5))).astype(F)


def hash32(i, j, frame):
    m = 0xFFFFFFFF
    h = ((i * 73856093) & m) ^ ((j * 19349663) & m) ^ ((frame * 83492791) & m)
    return h - (1 << 32) if h >= 1 << 31 else h


def brute_gabriel(pos, pairs_of):
    """Degrees from the Gabriel test over the given (i, j) pairs, searching
    every particle (fp32, same operation order and epsilon as the kernel)."""
    n = len(pos)
    deg = np.zeros(n, np.int64)
    for i, j in pairs_of:
        d = gw(pos[j] - pos[i])
        d2 = F(d[0] * d[0] + d[1] * d[1]) + F(d[2] * d[2])
        if d2 < F(1e-16):
            continue
        m = gw(pos[i] + F(0.5) * d)
        r_sq = F(0.25) * d2
        eps = max(F(1e-12), F(1e-6) * r_sq)
        e = gw(m[None, :] - pos)
        dk = (e[:, 0] * e[:, 0] + e[:, 1] * e[:, 1]) + e[:, 2] * e[:, 2]
        dk[[i, j]] = np.inf
        if not (dk < r_sq - eps).any():
            deg[i] += 1
            deg[j] += 1
    return deg


def cand_pairs(pairs, cnt):
    return [(i, int(pairs[i, p])) for i in range(len(cnt)) for p in range(cnt[i])]


def test_gabriel_collinear_middle_blocks(orc):
    pos = np.array([[-0.05, 0, 0], [0, 0, 0], [0.05, 0, 0]], F)
    rad = np.full(3, 0.03, F)
    G = 8
    p, cc, cs, ci = orc.grid_build(pos, 3, G, 1.0)
    pairs, cnt, _ = orc.topo_candidates(p, rad, 3, G, 1.0, F(2.2), 28, cs, cc, ci)
    deg = orc.topo_gabriel(p, 3, G, 1.0, 1.0 / G, cs, cc, ci, pairs, cnt, 28, 0, stride=1,
                           early_exit=False)
    assert list(deg) == [1, 2, 1]


@pytest.mark.parametrize("stride", [1, 3])
def test_gabriel_matches_brute_force(orc, stride):
    n, G = 400, 10
    pos, rad = fi.uniform(n, 0.005, 0.02, seed=3)
    p, cc, cs, ci = orc.grid_build(pos, n, G, 1.0)
    pairs, cnt, tr = orc.topo_candidates(p, rad, n, G, 1.0, F(2.2), 64, cs, cc, ci)
    assert not tr.any()
    deg = orc.topo_gabriel(p, n, G, 1.0, 1.0 / G, cs, cc, ci, pairs, cnt, 64, 5, stride=stride,
                           early_exit=False)
    keep = [(i, j) for i, j in cand_pairs(pairs, cnt)
            if stride == 1 or hash32(i, j, 5) % stride == 0]
    assert np.array_equal(deg, brute_gabriel(p, keep))
    # early exit skips pairs whose endpoints are both at or above the high degree
    ema = np.full(n, 20.0, F)
    assert not orc.topo_gabriel(p, n, G, 1.0, 1.0 / G, cs, cc, ci, pairs, cnt, 64, 5, stride,
                                True, 18.0, ema).any()


def test_batched_union_equals_full(orc):
    n, G, B, q = 400, 10, 5, 3
    pos, rad = fi.uniform(n, 0.005, 0.02, seed=4)
    p, cc, cs, ci = orc.grid_build(pos, n, G, 1.0)
    pairs, cnt, _ = orc.topo_candidates(p, rad, n, G, 1.0, F(2.2), 64, cs, cc, ci)
    full = orc.topo_gabriel(p, n, G, 1.0, 1.0 / G, cs, cc, ci, pairs, cnt, 64, 9, stride=q,
                            early_exit=False)
    deg = np.zeros(n, np.int32)
    for b in range(B):
        _, _, _, deg = orc.topo_gabriel_batched(p, rad, n, G, 1.0, 1.0 / G, F(2.2), 64, cs, cc,
                                                ci, b, B, q, 9, deg)
    assert np.array_equal(deg, full)


def test_knn_matches_brute_force(orc):
    n, G, k = 600, 8, 5
    pos, rad = fi.uniform(n, 0.01, 0.03, seed=5)
    p, cc, cs, ci = orc.grid_build(pos, n, G, 1.0)
    deg = orc.topo_knn(p, rad, n, G, 1.0, F(2.2), k, cs, cc, ci)
    for i in range(n):
        e = gw(p - p[i])
        d2 = (e[:, 0] * e[:, 0] + e[:, 1] * e[:, 1]) + e[:, 2] * e[:, 2]
        r = F(2.2) * rad[i]
        c = int(((d2 <= r * r) & (np.arange(n) != i)).sum())
        assert deg[i] == min(c, k)


def test_topo_ema_and_radius_rule(orc):
    deg = np.array([0, 10, 20], np.int32)
    ema = orc.topo_ema(deg, np.array([5.0, 5.0, 5.0], F), 3, F(0.1))
    assert np.allclose(ema, [4.5, 5.5, 6.5], rtol=1e-6)
    rad = orc.topo_radii(np.array([0.01, 0.01, 0.01, 0.0395], F), np.array([9, 14, 19, 3], F), 4,
                         10.0, 18.0, F(0.05), F(0.05), F(0.005), F(0.04))
    assert np.allclose(rad, [0.0105, 0.01, 0.0095, 0.04], rtol=1e-6)
