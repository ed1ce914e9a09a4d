"""GPU parity: photon-map build, radius query (KdTree::within) and gather
(estimate_radiance) against the FP64 oracle.

The index-set contract: refine64=1 must reproduce the reference's FP64
predicate ((0+dx^2)+dy^2)+dz^2 <= r2 bit for bit; refine64=0 must reproduce the
float predicate with the same operation order bit for bit (CPU brute force
in float).  Radiance must match the FP64 oracle within 1e-5 relative.
"""
import numpy as np
import pytest

import oracle_lib as O
from gpu_common import ctx, random_photons, say

import ppmpa_amd as P

pytestmark = pytest.mark.gpu


def _upload_and_build(pos, dir, wl, power=1.0):
    c = ctx()
    ph = P.PhotonSet(c, max(1, len(wl)))
    ph.upload(pos, dir, wl)
    m = P.PhotonMap(c, ph.capacity).build(ph, power)
    return m, ph


@pytest.mark.parametrize("n,nq,r,clustered,seed", [
    (1, 7, 0.3, False, 1),
    (100, 50, 0.2, False, 2),
    (5000, 400, 0.05, False, 3),
    (20000, 500, 0.02, True, 4),
    (64, 64, 0.5, False, 5),       # exactly one leaf
    (65, 64, 0.5, False, 6),       # one photon in the second leaf
    (4096 + 3, 300, 0.04, False, 7),
])
def test_radius_query_bitexact(n, nq, r, clustered, seed):
    say("radius_query", n, nq, r, clustered, seed)
    pos, dir, wl = random_photons(n, seed, clustered=clustered)
    rng = np.random.default_rng(seed + 100)
    # queries: random points plus points exactly on photons
    q = np.concatenate([rng.uniform(-0.05, 1.05, (nq - nq // 4, 3)), pos[rng.integers(0, n, nq // 4)]])
    m, _ = _upload_and_build(pos, dir, wl)
    r2 = r * r
    got32 = m.within(q, r2, refine64=False)
    want32 = O.within_brute(pos, q, r2, mode=1)
    for i in range(len(q)):
        assert np.array_equal(got32[i], want32[i]), f"float set differs at query {i}"
    got64 = m.within(q, r2, refine64=True)
    want64 = O.within_brute(pos, q, r2, mode=0)
    for i in range(len(q)):
        assert np.array_equal(got64[i], want64[i]), f"double set differs at query {i}"


def test_radius_query_boundary_photons():
    """Photons placed at float-rounding distance from the sphere: the double
    refinement must still match the FP64 predicate exactly."""
    say("radius_query boundary")
    rng = np.random.default_rng(11)
    q = np.array([[0.3, 0.6, 0.45]])
    r = 0.0123
    n = 3000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    eps = rng.uniform(-3e-7, 3e-7, n)
    pos = q + d * (r * (1.0 + eps))[:, None]
    wl = rng.integers(0, 3, n).astype(np.uint8)
    m, _ = _upload_and_build(pos, d, wl)
    got = m.within(q, r * r, refine64=True)[0]
    want = O.within_brute(pos, q, r * r, mode=0)[0]
    assert 0 < len(want) < n
    assert np.array_equal(got, want)
    got32 = m.within(q, r * r, refine64=False)[0]
    want32 = O.within_brute(pos, q, r * r, mode=1)[0]
    assert np.array_equal(got32, want32)


def test_kdtreetest_fixture():
    """kdtreetest.rs:19-50: six photons; within(origin, 2.5) and
    within(origin, 3.5^2)."""
    say("kdtreetest fixture")
    pos = np.array([[0, 0, 0], [1, 1, 0], [1, 1, 1], [-1, -1, -1], [2, 2, 2], [2, 2, 2]], float)
    dir = np.array([[0, 1, 0], [1, 0, 0], [0, 0, 1], [0, 1, 0], [1, 0, 0], [0, 0, 1]], float)
    wl = np.array([0, 1, 2, 0, 1, 2], np.uint8)
    m, _ = _upload_and_build(pos, dir, wl)
    assert list(m.within([[0, 0, 0]], 2.5, refine64=True)[0]) == [0, 1]
    assert list(m.within([[0, 0, 0]], 3.5 * 3.5, refine64=True)[0]) == [0, 1, 2, 3, 4, 5]


@pytest.mark.parametrize("pf", [0, 1, 2])
def test_gather_matches_oracle(pf):
    say("gather", pf)
    pos, dir, wl = random_photons(30000, 21, clustered=True)
    power = 1.0 / 30000
    m, _ = _upload_and_build(pos, dir, wl, power)
    om = O.OMap(pos, dir, wl, power)
    rng = np.random.default_rng(22)
    q = pos[rng.integers(0, len(wl), 2000)] + rng.normal(0, 0.003, (2000, 3))
    nrm = rng.normal(size=(2000, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    r2 = 0.01 ** 2
    got = m.estimate_radiance(q, nrm, r2, pfilter=pf, refine64=True)
    want = om.gather(q, nrm, r2, pf)
    scale = np.maximum(np.abs(want), 1e-300)
    nz = np.abs(want) > 0
    assert nz.sum() > 1000
    rel = np.abs(got - want)[nz] / scale[nz]
    assert rel.max() < 1e-5, rel.max()
    assert np.all(got[~nz] == 0)


def test_gather_uniform_carpet_closed_form():
    """A uniform carpet of density sigma on a plane hit at normal incidence:
    L = (1/(pi r^2)) * Σ P * 1 ~= sigma * P (the pi r^2 cancels)."""
    say("gather carpet")
    n_side = 400
    xs = (np.arange(n_side) + 0.5) / n_side
    X, Z = np.meshgrid(xs, xs)
    pos = np.stack([X.ravel(), np.zeros(X.size), Z.ravel()], 1)
    dir = np.tile([0.0, -1.0, 0.0], (pos.shape[0], 1))
    wl = np.zeros(pos.shape[0], np.uint8)
    power = 1e-3
    m, _ = _upload_and_build(pos, dir, wl, power)
    q = np.array([[0.5, 0.0, 0.5], [0.3, 0.0, 0.7]])
    nrm = np.array([[0.0, 1.0, 0.0]] * 2)
    r = 0.05
    got = m.estimate_radiance(q, nrm, r * r)
    sigma = n_side * n_side  # photons per m^2
    expect = sigma * power
    assert np.allclose(got[:, 0], expect, rtol=0.02), (got[:, 0], expect)
    assert np.all(got[:, 1:] == 0)


def test_empty_map():
    say("empty map")
    c = ctx()
    ph = P.PhotonSet(c, 8)
    ph.upload(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, np.uint8))
    m = P.PhotonMap(c, 8).build(ph, 1.0)
    assert [len(x) for x in m.within([[0, 0, 0], [1, 1, 1]], 1.0)] == [0, 0]
    assert np.all(m.estimate_radiance([[0, 0, 0]], [[0, 1, 0]], 1.0) == 0)


def test_nonfinite_photon_rejected():
    say("nonfinite")
    c = ctx()
    ph = P.PhotonSet(c, 8)
    with pytest.raises(P.PpmError) as e:
        ph.upload(np.array([[0, np.nan, 0]]), np.array([[0, 1, 0]]), np.array([0], np.uint8))
    assert e.value.status == -2


def test_golden_map_bitexact_sets_and_radiance():
    """Photon set dumped in the reference's text format (tests/golden):
    read_map -> device map; FP64-decided index sets equal the committed
    FP64 sets bit for bit; radiance within 1e-5 relative for every filter."""
    import os
    from ppmpa_amd import photonmap_io as M
    say("golden map")
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    z = np.load(os.path.join(gold, "queries_two_spheres.npz"), allow_pickle=False)
    c = ctx()
    m, ph = P.read_map(c, os.path.join(gold, "map_two_spheres.txt"))
    pm = M.parse_map(os.path.join(gold, "map_two_spheres.txt"))
    for j, r2 in enumerate(z["r2"]):
        got = m.within(z["q"], r2, refine64=True)
        off, idx = z[f"off_{j}"], z[f"idx_{j}"]
        for i in range(len(z["q"])):
            assert np.array_equal(got[i], idx[int(off[i]):int(off[i + 1])]), (j, i)
        got32 = m.within(z["q"], r2, refine64=False)
        want32 = O.within_brute(pm.pos, z["q"], r2, mode=1)
        for i in range(len(z["q"])):
            assert np.array_equal(got32[i], want32[i])
        for pf in (0, 1, 2):
            rad = m.estimate_radiance(z["q"], z["nrm"], r2, pfilter=pf, refine64=True)
            want = z[f"rad_{j}_{pf}"]
            nz = want != 0
            assert np.all(rad[~nz] == 0)
            assert np.max(np.abs(rad[nz] - want[nz]) / np.abs(want[nz])) < 1e-5


def test_float_gather_production_arithmetic():
    """The production gather decides membership in float (no refinement):
    it matches the FP64 oracle within 1e-5 except for queries with a photon
    within float rounding of the sphere (a boundary flip changes the sum by
    one photon)."""
    say("float gather")
    pos, dir, wl = random_photons(30000, 31, clustered=True)
    m, _ = _upload_and_build(pos, dir, wl, 1e-4)
    om = O.OMap(pos, dir, wl, 1e-4)
    rng = np.random.default_rng(32)
    q = pos[rng.integers(0, len(wl), 3000)] + rng.normal(0, 0.003, (3000, 3))
    nrm = rng.normal(size=(3000, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    got = m.estimate_radiance(q, nrm, 0.01 ** 2, refine64=False)
    want = om.gather(q, nrm, 0.01 ** 2, 0)
    rel = np.abs(got - want).max(1) / np.maximum(np.abs(want).max(1), 1e-300)
    assert (rel < 1e-5).mean() > 0.99
