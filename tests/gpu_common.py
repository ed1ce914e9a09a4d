"""Shared helpers of the GPU parity tests."""
import os
import sys

import numpy as np
import pytest

import ppmpa_amd as P
from ppmpa_amd import _abi as A

_CTX = None


def ctx():
    """One context for the whole test process (one process uses the GPU)."""
    global _CTX
    if _CTX is None:
        if not os.path.exists(A.LIB_PATH):
            pytest.fail("libppm.so is not built; run python -m ppmpa_amd.build")
        _CTX = P.Context(0)
    return _CTX


def say(*a):
    """Name the case before a GPU step, flushed, so a crash log shows it."""
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def small_scene(kind="two_spheres", xreso=64, yreso=48):
    if kind == "two_spheres":
        return P.two_spheres_scene(xreso, yreso)
    return P.caustic_scene(xreso, yreso)


def random_photons(n, seed, lo=0.0, hi=1.0, clustered=False):
    rng = np.random.default_rng(seed)
    if clustered:
        centers = rng.uniform(lo, hi, (max(1, n // 200), 3))
        pos = centers[rng.integers(0, centers.shape[0], n)] + rng.normal(0, 0.01, (n, 3))
    else:
        pos = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    wl = rng.integers(0, 3, n).astype(np.uint8)
    return pos, d, wl
