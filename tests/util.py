"""Shared input builders for the tests (seeded, reference layout, x fastest)."""
import numpy as np


def faces(rng, W, H, D, scale=1.0):
    u = (rng.standard_normal((D, H, W + 1)) * scale).astype(np.float32)
    v = (rng.standard_normal((D, H + 1, W)) * scale).astype(np.float32)
    w = (rng.standard_normal((D + 1, H, W)) * scale).astype(np.float32)
    return u, v, w


def occupancy(rng, W, H, D, frac=0.1, shell=True):
    occ = (rng.random((D, H, W)) < frac).astype(np.uint8)
    if shell:
        occ[0] = occ[-1] = 1
        occ[:, 0] = occ[:, -1] = 1
        occ[:, :, 0] = occ[:, :, -1] = 1
    return occ


def cells(rng, W, H, D, lo=0.0, hi=1.0):
    return rng.uniform(lo, hi, (D, H, W)).astype(np.float32)


def copies(*arrs):
    return [a.copy() for a in arrs]
