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


def keyed_photons(meta, pos, dir, wl):
    """Photons ordered by (photon id, depth), the key both sides share."""
    pid = meta >> 7
    depth = (meta >> 2) & 31
    keys = pid.astype(np.int64) * 64 + depth
    order = np.argsort(keys)
    return keys[order], pos[order], dir[order], wl[order]


def check_photon_trace_crn(c, sc, seed, pass_, n, uc, maxd, match=0.99):
    """Device photon walks vs the FP64 oracle's on the same Philox streams:
    the same (photon, depth) records except where a float decision flipped,
    equal channels, positions and directions within 1e-4 at the 99th
    percentile.  Returns the device photon set."""
    import oracle_lib as O
    ds = P.DeviceScene(c, sc)
    ph = P.trace_photons(c, ds, seed=seed, pass_=pass_, n_photon=n, use_classic=uc, max_depth=maxd)
    gpos, gdir, gwl, gmeta = ph.download()
    osc = O.OScene(sc)
    opos, odir, owl, ometa, opw = osc.trace_photons(seed, pass_, n, uc, maxd)
    assert abs(opw - ph.power) <= 1e-15 * opw
    gk, gpos, gdir, gwl = keyed_photons(gmeta, gpos, gdir, gwl)
    ok, opos, odir, owl = keyed_photons(ometa, opos, odir, owl)
    assert len(np.unique(gk)) == len(gk)
    assert len(ok) > 0
    common, gi, oi = np.intersect1d(gk, ok, return_indices=True)
    assert len(common) >= match * len(ok), (len(common), len(ok), len(gk))
    assert abs(len(gk) - len(ok)) <= (1.0 - match) * len(ok) + 1
    assert np.all(gwl[gi] == owl[oi])
    dpos = np.abs(gpos[gi] - opos[oi]).max(axis=1)
    assert np.quantile(dpos, 0.99) < 1e-4, np.quantile(dpos, 0.99)
    ddir = np.abs(gdir[gi] - odir[oi]).max(axis=1)
    assert np.quantile(ddir, 0.99) < 1e-4, np.quantile(ddir, 0.99)
    return ph


def eye_agreement(got, want, tol=1e-3):
    """Fraction of pixels whose channels all agree within tol relative
    (floor: tol of the image maximum)."""
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-3 * np.abs(want).max())
    return float((rel.max(axis=2) < tol).mean())


def check_eye_crn(c, sc, seed, n, r2, uc, maxd, pfilter=0, frac_min=0.98):
    """One eye pass over the oracle's photons: device vs FP64 oracle."""
    import oracle_lib as O
    osc = O.OScene(sc)
    pos, dir, wl, meta, pw = osc.trace_photons(seed, 1, n, uc, maxd)
    ph = P.PhotonSet(c, max(1, len(wl)))
    ph.upload(pos, dir, wl)
    ds = P.DeviceScene(c, sc)
    m = P.PhotonMap(c, ph.capacity).build(ph, pw, ds)
    got = P.trace_eye(c, ds, m, seed=seed + 1, pass_=1, r2=r2, use_classic=uc, max_depth=maxd, pfilter=pfilter)
    om = O.OMap(pos, dir, wl, pw)
    want = osc.trace_eye(om, seed + 1, 1, r2, uc, maxd, pfilter)
    assert np.isfinite(got).all()
    assert want.max() > 0
    frac = eye_agreement(got, want)
    assert frac > frac_min, frac
    assert abs(got.mean() - want.mean()) < 1e-3 * want.mean()
    return got, want
