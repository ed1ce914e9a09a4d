"""ppm_gather_noalloc, the gather that allocates nothing, against ppm_gather
through the C ABI: bit-for-bit equal results on the tree and the grid map
forms, for every photon filter, on two calls in a row (the second with fewer
queries, so a stale result from the first would show), at the bound, and
PPM_E_RANGE beyond it."""
import ctypes as C

import numpy as np
import pytest

from gpu_common import ctx, random_photons, say

import ppmpa_amd as P
from ppmpa_amd import _abi as A

pytestmark = pytest.mark.gpu

MAXQ = 65536  # PPM_NOALLOC_MAX_QUERIES


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _gather(c, fn, m, q, nrm, r2, pf, refine):
    q = np.ascontiguousarray(q, np.float64).reshape(-1, 3)
    nrm = np.ascontiguousarray(nrm, np.float64).reshape(-1, 3)
    out = np.full(q.shape, np.nan)
    s = getattr(c.lib, fn)(c.h, m.h, _dp(q), _dp(nrm), q.shape[0], float(r2), int(pf), int(refine), _dp(out))
    return s, out


def _maps(n=30000, seed=61):
    c = ctx()
    pos, dir, wl = random_photons(n, seed, clustered=True)
    ph = P.PhotonSet(c, n)
    ph.upload(pos, dir, wl)
    r2 = 0.012 ** 2
    tree = P.PhotonMap(c, n).build(ph, 1.0 / n)
    grid = P.PhotonMap(c, n).build_grid(ph, 1.0 / n, r2)
    return c, pos, ph, tree, grid, r2


def _queries(pos, nq, seed):
    rng = np.random.default_rng(seed)
    q = pos[rng.integers(0, len(pos), nq)] + rng.normal(0, 0.004, (nq, 3))
    nrm = rng.normal(size=(nq, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    return q, nrm


@pytest.mark.parametrize("form", ["tree", "grid"])
def test_noalloc_equals_gather_two_calls(form):
    c, pos, _ph, tree, grid, r2 = _maps()
    m = tree if form == "tree" else grid
    refines = (0, 1) if form == "tree" else (0,)
    for call, nq in enumerate((6000, 700)):
        q, nrm = _queries(pos, nq, 100 + call)
        for pf in (0, 1, 2):
            for refine in refines:
                say("noalloc", form, "call", call, "nq", nq, "filter", pf, "refine", refine)
                s0, want = _gather(c, "ppm_gather", m, q, nrm, r2, pf, refine)
                assert s0 == 0, c.lib.ppm_last_error(c.h)
                s1, got = _gather(c, "ppm_gather_noalloc", m, q, nrm, r2, pf, refine)
                assert s1 == 0, c.lib.ppm_last_error(c.h)
                assert (want != 0).sum() > nq  # a real gather, not empty balls
                assert np.array_equal(got, want), (form, call, pf, refine, np.abs(got - want).max())


def test_noalloc_bound():
    c, pos, _ph, tree, grid, r2 = _maps(8000, 62)
    q, nrm = _queries(pos, MAXQ + 1, 7)
    for m in (tree, grid):
        say("noalloc bound", m is grid)
        s, _ = _gather(c, "ppm_gather_noalloc", m, q, nrm, r2, 0, 0)
        assert s == A.PPM_E_RANGE
        s0, want = _gather(c, "ppm_gather", m, q[:MAXQ], nrm[:MAXQ], r2, 0, 0)
        s1, got = _gather(c, "ppm_gather_noalloc", m, q[:MAXQ], nrm[:MAXQ], r2, 0, 0)
        assert s0 == 0 and s1 == 0
        assert np.array_equal(got, want)
