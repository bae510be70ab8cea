"""Shared construction of multi-species test cases, after the reference's fused-contact test
(apps/CCOptimization/tests/test_fused_contact.cpp:97-160): cells of side N^(1/d) (optionally skewed), radii from
a packing fraction (equal, or spread over 0.7-1.3 times the mean), species labels interleaved by a seeded shuffle,
and each species' k-set at its own number density and chi with weights 1 + 0.1 s."""
import math

import numpy as np

from oracle import pyoracle as po


def cell(n, d, skew=False):
    L = n ** (1.0 / d)
    a = np.eye(d) * L
    if skew and d >= 2:
        a[1, 0] = 0.27 * L
    if skew and d == 3:
        a[2, 0], a[2, 1] = -0.19 * L, 0.21 * L
    return a


def ball_volume(d, r):
    return math.pi ** (d / 2) / math.gamma(d / 2 + 1) * r ** d


def radii_of(n, d, phi, unequal):
    mean = (phi / ball_volume(d, 1.0)) ** (1.0 / d)
    r = np.full(n, mean)
    if unequal:
        r = mean * (0.7 + 0.6 * np.fmod(0.6180339887 * np.arange(n), 1.0))
    return r


def species_setup(a, sizes, chi, seed=13):
    """Interleaved species labels (a seeded shuffle) and each species' k-set at its own density and chi."""
    d = a.shape[0]
    n = sum(sizes)
    label = np.concatenate([np.full(m, q) for q, m in enumerate(sizes)])
    rng = np.random.default_rng(seed)
    rng.shuffle(label)
    idx = [list(np.nonzero(label == q)[0]) for q in range(len(sizes))]
    ks, vs = [], []
    V = po.volume(a)
    for q, m in enumerate(sizes):
        rho_s = m / V
        K = 2 * math.pi * (2 * d * rho_s * chi / ball_volume(d, 1.0)) ** (1.0 / d)
        k, _ = po.get_ks(a, K)
        ks.append(k)
        vs.append(np.full(len(k), 1.0 + 0.1 * q))
    return idx, ks, vs


