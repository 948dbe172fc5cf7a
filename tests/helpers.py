"""Shared helpers of the test suite."""
import numpy as np


def unit(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v, axis=-1, keepdims=True)).astype(np.float32)


def random_dirs(rng, n):
    return unit(rng.normal(size=(n, 3)))


def hemi_dirs(rng, n, nrm):
    d = random_dirs(rng, n)
    s = np.sign(np.sum(d * nrm, axis=1, keepdims=True))
    s[s == 0] = 1
    return (d * s).astype(np.float32)


def rel_err(a, b, floor=1e-6):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), floor)
