"""Synthetic clustered int8 data: a Gaussian mixture rounded and clipped to int8.

Recipe: centers iid U[-60, 60]^dim; each point picks a component uniformly and adds
iid N(0, 20^2) per dimension; round half to even (numpy.rint); clip to [-128, 127].
Centers, base points and queries come from separate PCG64 streams (default seeds 1, 2,
3), so queries are held-out draws from the same mixture.
"""
import hashlib

import numpy as np


def gmm_centers(ncomp, dim, seed=1, center_range=60.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.uniform(-center_range, center_range, size=(ncomp, dim))


def gmm_points(centers, n, seed, sigma=20.0, chunk=1 << 18):
    rng = np.random.Generator(np.random.PCG64(seed))
    ncomp, dim = centers.shape
    out = np.empty((n, dim), dtype=np.int8)
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        comp = rng.integers(0, ncomp, size=m)
        x = centers[comp] + rng.normal(0.0, sigma, size=(m, dim))
        np.clip(np.rint(x), -128, 127, out=x)
        out[s:s + m] = x.astype(np.int8)
    return out


def gmm_dataset(n, nq, dim=100, ncomp=32, seeds=(1, 2, 3), sigma=20.0, center_range=60.0):
    """Returns (base [n, dim] int8, queries [nq, dim] int8)."""
    c = gmm_centers(ncomp, dim, seeds[0], center_range)
    return gmm_points(c, n, seeds[1], sigma), gmm_points(c, nq, seeds[2], sigma)


def sha256(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
