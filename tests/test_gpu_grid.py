"""GPU parity of the passes' own gather: the hashed photon grid that every
PPM-PA pass builds (k_grid_* kernels) and the shade kernel that gathers over
it (k_shade), run on uploaded photon sets through ppm_map_build_grid.

Contracts (KdTree::within, tracer.rs:180; estimate_radiance,
tracer.rs:179-195):
  - index sets through the shade kernel's cell walk equal the float brute
    force ((dx^2 + dy^2) + dz^2 <= r2 in float, no contraction) bit for bit,
    and, refined in double, the FP64 brute force bit for bit;
  - the shade kernel's radiance is within 1e-5 relative of the FP64 oracle's
    estimate over the same (float-decided) ball, for all three filters, and
    exactly zero where the oracle's is.
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from gpu_common import ctx, random_photons, say

import ppmpa_amd as P
from ppmpa_amd import _abi as A

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RTOL = 1e-5


def _grid(pos, dir, wl, power, r2, dscene=None):
    c = ctx()
    ph = P.PhotonSet(c, max(1, len(wl)))
    ph.upload(pos, dir, wl)
    m = P.PhotonMap(c, ph.capacity).build_grid(ph, power, r2, dscene)
    return m, ph


def _check_sets(m, pos, q, r2):
    got32 = m.within(q, r2, refine64=False)
    want32 = O.within_brute(pos, q, r2, mode=1)
    for i in range(len(q)):
        assert np.array_equal(got32[i], want32[i]), f"float set differs at query {i}"
    got64 = m.within(q, r2, refine64=True)
    want64 = O.within_brute(pos, q, r2, mode=0)
    for i in range(len(q)):
        assert np.array_equal(got64[i], want64[i]), f"double set differs at query {i}"
    return want32


# Largest filter weight (tracer.rs:197-216): cone 1/(1 - 2/(3*1.1)), Gauss 0.918 + 0.5.
WMAX = {0: 1.0, 1: 1.0 / (1.0 - 2.0 / 3.3), 2: 1.418}
# Float rounding allowances per photon term, absolute (so a sum of a few
# small terms can carry a large relative error):
#  - the cosine -n.dir from the float normal and direction, three products
#    and two sums: below 4e-7, taken as 1e-6 (times the filter's weight);
#  - the filter's weight (cone, Gauss) of d2/r2 formed from float
#    coordinates: p - q cancels when |p| >> r, so |d|/r carries up to
#    sqrt(3) * 2^-24 * (2|q| + 2r) / r and the weight, whose slope in |d|/r is
#    at most 2.3 (cone) and 2.1 (Gauss), up to 8 * 2^-24 * (|q| + r) / r,
#    taken as 1e-6 * (|q| + r) / r.
COS_ABS = 1e-6
T_ABS = 1e-6


def _check_radiance(m, om, q, nrm, r2, pfs=(0, 1, 2), min_nonzero=1):
    """Per channel, |got - want| <= 1e-5 |want| + (P / (pi r2)) * n_ch *
    (COS_ABS * wmax + T_ABS * (|q| + r) / r for a filter), n_ch the channel's
    photons in the ball; with no filter, at least 99% of the non-zero
    entries within 1e-5 relative alone."""
    q = np.asarray(q, float).reshape(-1, 3)
    sets = O.within_brute(om.pos, q, r2, mode=1)
    counts = np.array([np.bincount(om.wl[s], minlength=3) for s in sets], float).reshape(-1, 3)
    unit = om.power / (np.pi * r2)
    r = np.sqrt(r2)
    qmag = (np.abs(q).max(axis=1) + r) / r
    for pf in pfs:
        got = m.estimate_radiance(q, nrm, r2, pfilter=pf)
        want = om.gather(q, nrm, r2, pf, f32set=True)
        nz = want != 0
        assert nz.sum() >= min_nonzero
        assert np.all(got[~nz] == 0), f"filter {pf}: non-zero where the oracle is zero"
        err = np.abs(got - want)
        per_term = COS_ABS * WMAX[pf] + (T_ABS * qmag[:, None] if pf else 0.0)
        bound = RTOL * np.abs(want) + unit * counts * per_term
        assert np.all(err <= bound), (pf, np.max(err / np.maximum(bound, 1e-300)))
        if pf == 0 and nz.any():
            rel = err[nz] / np.abs(want[nz])
            assert (rel < RTOL).mean() >= 0.99, (pf, (rel < RTOL).mean())


@pytest.mark.parametrize("n,nq,r,clustered,seed", [
    (1, 7, 0.3, False, 1),
    (100, 50, 0.2, False, 2),
    (5000, 400, 0.05, False, 3),
    (20000, 500, 0.02, True, 4),
    (4096 + 3, 300, 0.04, False, 7),
    # about 300 tiles of the sort: its look-back over many earlier tiles
    (1_200_000, 200, 0.01, True, 11),
])
def test_grid_sets_and_radiance(n, nq, r, clustered, seed):
    say("grid sets+radiance", n, nq, r, clustered, seed)
    pos, dir, wl = random_photons(n, seed, clustered=clustered)
    rng = np.random.default_rng(seed + 100)
    q = np.concatenate([rng.uniform(-0.05, 1.05, (nq - nq // 4, 3)), pos[rng.integers(0, n, nq // 4)]])
    nrm = rng.normal(size=q.shape)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    r2 = r * r
    power = 1.0 / max(n, 1)
    m, _ = _grid(pos, dir, wl, power, r2)
    sets = _check_sets(m, pos, q, r2)
    om = O.OMap(pos, dir, wl, power)
    _check_radiance(m, om, q, nrm, r2, min_nonzero=0 if n < 1000 else 50)
    # a smaller radius on the same grid (the cells still cover the ball)
    _check_sets(m, pos, q, 0.37 * r2)
    _check_radiance(m, om, q, nrm, 0.37 * r2, pfs=(0,), min_nonzero=0)
    assert sum(len(s) for s in sets) > 0


def test_grid_boundary_photons():
    """Photons at float-rounding distance from the sphere: the float sets
    are the float predicate's, the refined ones the FP64 predicate's, and
    the radiance is the oracle's over the float ball."""
    say("grid boundary")
    rng = np.random.default_rng(11)
    q = np.array([[0.3, 0.6, 0.45]])
    r = 0.0123
    n = 3000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pos = q + d * (r * (1.0 + rng.uniform(-3e-7, 3e-7, n)))[:, None]
    wl = rng.integers(0, 3, n).astype(np.uint8)
    m, _ = _grid(pos, d, wl, 1e-3, r * r)
    want64 = O.within_brute(pos, q, r * r, mode=0)[0]
    want32 = O.within_brute(pos, q, r * r, mode=1)[0]
    assert 0 < len(want64) < n and not np.array_equal(want64, want32)
    _check_sets(m, pos, q, r * r)
    nrm = -d[:1]
    _check_radiance(m, O.OMap(pos, d, wl, 1e-3), q, nrm, r * r)


def test_grid_kdtreetest_fixture():
    """kdtreetest.rs:19-50: within(origin, 2.5) = {p0, p1};
    within(origin, 3.5^2) = all six."""
    say("grid kdtreetest")
    pos = np.array([[0, 0, 0], [1, 1, 0], [1, 1, 1], [-1, -1, -1], [2, 2, 2], [2, 2, 2]], float)
    dir = np.array([[0, 1, 0], [1, 0, 0], [0, 0, 1], [0, 1, 0], [1, 0, 0], [0, 0, 1]], float)
    wl = np.array([0, 1, 2, 0, 1, 2], np.uint8)
    m, _ = _grid(pos, dir, wl, 1.0, 3.5 * 3.5)
    assert list(m.within([[0, 0, 0]], 2.5, refine64=True)[0]) == [0, 1]
    assert list(m.within([[0, 0, 0]], 2.5)[0]) == [0, 1]
    assert list(m.within([[0, 0, 0]], 3.5 * 3.5, refine64=True)[0]) == [0, 1, 2, 3, 4, 5]


def test_grid_golden_map():
    """The committed photon set in the reference's text format: FP64-decided
    sets equal the committed FP64 sets, float sets the float brute force, and
    the shade kernel's radiance is the oracle's over the float ball for every
    filter (and the committed FP64 radiance, no filter, wherever the two
    balls agree), to the float bounds of _check_radiance."""
    from ppmpa_amd import photonmap_io as M
    say("grid golden map")
    z = np.load(os.path.join(GOLD, "queries_two_spheres.npz"), allow_pickle=False)
    pm = M.parse_map(os.path.join(GOLD, "map_two_spheres.txt"))
    dn = pm.dir / np.linalg.norm(pm.dir, axis=1, keepdims=True)
    om = O.OMap(pm.pos, dn, pm.wl, pm.power)
    for j, r2 in enumerate(z["r2"]):
        m, _ = _grid(pm.pos, pm.dir, pm.wl, pm.power, r2)
        got = m.within(z["q"], r2, refine64=True)
        off, idx = z[f"off_{j}"], z[f"idx_{j}"]
        for i in range(len(z["q"])):
            assert np.array_equal(got[i], idx[int(off[i]):int(off[i + 1])]), (j, i)
        s32 = _check_sets(m, pm.pos, z["q"], r2)
        _check_radiance(m, om, z["q"], z["nrm"], r2, min_nonzero=100)
        # the committed FP64 radiance (no filter) wherever the balls agree
        same = np.array([np.array_equal(s32[i], idx[int(off[i]):int(off[i + 1])]) for i in range(len(z["q"]))])
        rad = m.estimate_radiance(z["q"], z["nrm"], r2, pfilter=0)
        want = z[f"rad_{j}_0"]
        nz = (want != 0) & same[:, None]
        assert nz.sum() > 100
        assert (np.abs(rad[nz] - want[nz]) / np.abs(want[nz]) < RTOL).mean() >= 0.99


def test_grid_scene_bounds_clamp():
    """Photons beyond the scene's bounds keep cells of their own (hashed like
    any other), out to about 2^20 cells, where the coordinates clamp: near
    and kilometres-away photons and queries, sets exact (a scene's bounds
    are what a pass's grid uses)."""
    say("grid clamp")
    c = ctx()
    ds = P.DeviceScene(c, P.two_spheres_scene(16, 16))
    rng = np.random.default_rng(41)
    n = 6000
    pos = rng.uniform(-0.6, 1.6, (n, 3))
    # a plane far out, as infinite walls store them: 1 km and 200 km away
    far = rng.uniform(-1.0, 1.0, (600, 3)) * np.array([40.0, 0.0, 40.0]) + np.array([0.5, 0.0, -1000.0])
    farther = rng.uniform(-1.0, 1.0, (300, 3)) * np.array([300.0, 0.0, 300.0]) + np.array([0.5, 0.0, -2e5])
    pos = np.concatenate([pos, far, farther])
    n = len(pos)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    wl = rng.integers(0, 3, n).astype(np.uint8)
    r2 = 0.08 ** 2
    m, _ = _grid(pos, d, wl, 1.0 / n, r2, ds)
    q = np.concatenate([rng.uniform(-0.7, 1.7, (300, 3)), pos[:100], far[:50], farther[:50]])
    nrm = rng.normal(size=q.shape)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    _check_sets(m, pos, q, r2)
    _check_radiance(m, O.OMap(pos, d, wl, 1.0 / n), q, nrm, r2, min_nonzero=100)


def test_grid_uniform_carpet_closed_form():
    """Uniform carpet at normal incidence, no filter: L = sigma * P."""
    say("grid carpet")
    n_side = 400
    xs = (np.arange(n_side) + 0.5) / n_side
    X, Z = np.meshgrid(xs, xs)
    pos = np.stack([X.ravel(), np.zeros(X.size), Z.ravel()], 1)
    dirs = np.tile([0.0, -1.0, 0.0], (pos.shape[0], 1))
    wl = np.zeros(pos.shape[0], np.uint8)
    power = 1e-3
    m, _ = _grid(pos, dirs, wl, power, 0.05 ** 2)
    got = m.estimate_radiance([[0.5, 0.0, 0.5], [0.3, 0.0, 0.7]], [[0.0, 1.0, 0.0]] * 2, 0.05 ** 2)
    assert np.allclose(got[:, 0], n_side * n_side * power, rtol=0.02), got[:, 0]
    assert np.all(got[:, 1:] == 0)


def test_grid_empty_and_errors():
    say("grid empty/errors")
    c = ctx()
    ph = P.PhotonSet(c, 8)
    ph.upload(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, np.uint8))
    m = P.PhotonMap(c, 8).build_grid(ph, 1.0, 0.25)
    assert [len(x) for x in m.within([[0, 0, 0], [1, 1, 1]], 0.25)] == [0, 0]
    assert np.all(m.estimate_radiance([[0, 0, 0]], [[0, 1, 0]], 0.25) == 0)
    with pytest.raises(P.PpmError) as e:
        m.within([[0, 0, 0]], 0.3)  # beyond the grid's radius
    assert e.value.status == A.PPM_E_INVALID
    with pytest.raises(P.PpmError) as e:
        m.estimate_radiance([[0, 0, 0]], [[0, 1, 0]], 0.25, refine64=True)
    assert e.value.status == A.PPM_E_INVALID
    with pytest.raises(P.PpmError):
        P.PhotonMap(c, 8).build_grid(ph, 1.0, -1.0)


def test_grid_eye_pass_matches_tree():
    """ppm_trace_eye over the grid of a traced photon set shades through the
    pass's kernel: every pixel equals the tree map's to float rounding."""
    say("grid eye pass")
    c = ctx()
    sc = P.two_spheres_scene(40, 32)
    ds = P.DeviceScene(c, sc)
    ph = P.trace_photons(c, ds, seed=3, pass_=1, n_photon=40000, max_depth=8)
    r2 = 0.02 ** 2
    mt = P.PhotonMap(c, ph.capacity).build(ph, ph.power, ds)
    mg = P.PhotonMap(c, ph.capacity).build_grid(ph, ph.power, r2, ds)
    a = P.trace_eye(c, ds, mt, seed=4, pass_=1, r2=r2, max_depth=8)
    b = P.trace_eye(c, ds, mg, seed=4, pass_=1, r2=r2, max_depth=8)
    assert a.max() > 0
    rel = np.abs(a - b) / np.maximum(np.abs(a), 1e-6 * a.max())
    assert np.quantile(rel, 0.999) < 1e-5 and rel.max() < 1e-3, rel.max()
