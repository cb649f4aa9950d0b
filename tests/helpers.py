"""Shared helpers of the parity tests."""
from __future__ import annotations

import math

import numpy as np


def random_rays(n, lo, hi, seed=1):
    rng = np.random.default_rng(seed)
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    o = lo + (hi - lo) * rng.random((n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d


def canonical_trace(facets, kinds):
    """Sort runs of consecutive transparent passes (kind 3) by facet id: the
    BVH walk order of passes between two hard events is not part of the
    physics."""
    f = facets.copy()
    k = kinds.copy()
    for row in range(f.shape[0]):
        j = 0
        n = f.shape[1]
        while j < n:
            if k[row, j] == 3:
                e = j
                while e < n and k[row, e] == 3:
                    e += 1
                f[row, j:e] = np.sort(f[row, j:e])
                j = e
            else:
                j += 1
    return f, k


def z_two_proportions(k1, n1, k2, n2):
    p = (k1 + k2) / (n1 + n2)
    s = math.sqrt(max(p * (1 - p) * (1 / n1 + 1 / n2), 1e-300))
    return (k1 / n1 - k2 / n2) / s


def z_means(m1, s1, m2, s2):
    return (m1 - m2) / math.sqrt(max(s1 ** 2 + s2 ** 2, 1e-300))
