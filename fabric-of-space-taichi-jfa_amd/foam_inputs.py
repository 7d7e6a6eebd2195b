"""Synthetic foam inputs (positions in [-L/2, L/2)^3, radii) with digests.

Positions are drawn in [0, L)^3 and wrapped once with p - L*floor(p/L + 1/2)
(the always-wrapped invariant of run.py:234-237), so every consumer sees the
same fp32 bits.
"""
import hashlib

import numpy as np


def wrap(p, L=1.0):
    p = np.asarray(p, np.float32)
    L = np.float32(L)
    return (p - L * np.floor(p / L + np.float32(0.5))).astype(np.float32)


def uniform(n, r_lo, r_hi, seed=0, L=1.0):
    """Uniform positions, uniform radii (rng = default_rng(seed))."""
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3), dtype=np.float32) * np.float32(L)
    rad = rng.uniform(r_lo, r_hi, n).astype(np.float32)
    return wrap(pos, L), rad


def clustered(n, seed=0, centres=64, sigma=0.05, median=0.006, sigma_log=0.3,
              r_lo=0.003, r_hi=0.015):
    """Half of the particles in Gaussian clusters, half uniform; lognormal radii."""
    rng = np.random.default_rng(seed)
    c = rng.random((centres, 3))
    half = n // 2
    which = rng.integers(0, centres, half)
    pc = c[which] + rng.normal(0.0, sigma, (half, 3))
    pu = rng.random((n - half, 3))
    pos = np.concatenate([pc, pu]).astype(np.float32)
    pos = pos - np.floor(pos)
    rad = np.exp(rng.normal(np.log(median), sigma_log, n))
    rad = np.clip(rad, r_lo, r_hi).astype(np.float32)
    return wrap(pos), rad


def for_config(cfg):
    if cfg.name in ("c1",):
        return uniform(cfg.n, 0.02, 0.04, seed=0)
    if cfg.name in ("c2", "c2c", "c2t"):
        return uniform(cfg.n, 0.012, 0.024, seed=0)
    if cfg.name == "c3":
        return clustered(cfg.n, seed=0)
    if cfg.name in ("c5", "c5s"):
        return uniform(cfg.n, 0.005, 0.011, seed=0)
    raise KeyError(cfg.name)


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()
