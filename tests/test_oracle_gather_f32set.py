"""The oracle's radiance estimate over the float-decided ball of the GPU
passes (orc_gather_f32set), pinned against the oracle's own FP64
estimate_radiance (tracer.rs:179-195) and the brute-force index sets.
CPU only."""
import os

import numpy as np

import oracle_lib as O
from ppmpa_amd import photonmap_io as M

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden():
    m = M.parse_map(os.path.join(GOLD, "map_two_spheres.txt"))
    z = np.load(os.path.join(GOLD, "queries_two_spheres.npz"), allow_pickle=False)
    dn = m.dir / np.linalg.norm(m.dir, axis=1, keepdims=True)
    return m, dn, z


def test_equal_sets_give_the_fp64_estimate_bit_for_bit():
    """Where the float and double balls hold the same photons, the estimate
    over the float ball is the FP64 estimate itself (same terms, same order)."""
    m, dn, z = _golden()
    om = O.OMap(m.pos, dn, m.wl, m.power)
    for r2 in z["r2"]:
        s64 = O.within_brute(m.pos, z["q"], r2, mode=0)
        s32 = O.within_brute(m.pos, z["q"], r2, mode=1)
        same = np.array([np.array_equal(a, b) for a, b in zip(s64, s32)])
        assert same.mean() > 0.99
        for pf in (0, 1, 2):
            a = om.gather(z["q"], z["nrm"], r2, pf)
            b = om.gather(z["q"], z["nrm"], r2, pf, f32set=True)
            assert np.array_equal(a[same], b[same])


def test_boundary_photon_changes_exactly_its_own_term():
    """A photon inside the float ball but outside the double one (a set built
    at float rounding distance from the sphere) adds exactly its own term to
    the float-ball estimate: the difference is P * max(0, -n.dir) / (pi r2)
    in its channel, and zero elsewhere."""
    q = np.array([0.3, 0.6, 0.45])
    r = 0.0123
    r2 = r * r
    rng = np.random.default_rng(5)
    d = rng.normal(size=(4000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pos = q + d * (r * (1.0 + rng.uniform(-3e-7, 3e-7, 4000)))[:, None]
    s64 = O.within_brute(pos, q, r2, mode=0)[0]
    s32 = O.within_brute(pos, q, r2, mode=1)[0]
    only32 = np.setdiff1d(s32, s64)
    only64 = np.setdiff1d(s64, s32)
    assert len(only32) + len(only64) > 0
    # keep one flipped photon and the photons both balls agree on, far apart
    keep = np.concatenate([only32[:1], only64[:1]]) if len(only32) and len(only64) else \
        (only32[:1] if len(only32) else only64[:1])
    both = np.intersect1d(s32, s64)[:20]
    sel = np.concatenate([keep, both])
    wl = np.zeros(len(sel), np.uint8)
    wl[0] = 1
    dirs = np.tile([0.0, -1.0, 0.0], (len(sel), 1))
    nrm = np.array([[0.0, 1.0, 0.0]])
    power = 0.25
    om = O.OMap(pos[sel], dirs, wl, power)
    a = om.gather(q, nrm, r2, 0)[0]
    b = om.gather(q, nrm, r2, 0, f32set=True)[0]
    term = power / (np.pi * r2)
    if len(only32) and len(only64):
        # first photon (channel 1) in the float ball only, the second
        # (channel 0) in the double ball only
        assert np.isclose(b[1] - a[1], term, rtol=1e-12)
        assert np.isclose(a[0] - b[0], term, rtol=1e-12)
    elif len(only32):
        assert np.isclose(b[1] - a[1], term, rtol=1e-12)
        assert b[0] == a[0]
    else:
        assert np.isclose(a[1] - b[1], term, rtol=1e-12)
        assert b[0] == a[0]
    assert a[2] == 0 and b[2] == 0


def test_carpet_closed_form():
    """Uniform carpet, normal incidence, no filter: L = sigma * P over either
    ball (the pi r^2 cancels up to the count's discreteness)."""
    n_side = 300
    xs = (np.arange(n_side) + 0.5) / n_side
    X, Z = np.meshgrid(xs, xs)
    pos = np.stack([X.ravel(), np.zeros(X.size), Z.ravel()], 1)
    dirs = np.tile([0.0, -1.0, 0.0], (pos.shape[0], 1))
    wl = np.full(pos.shape[0], 2, np.uint8)
    power = 2e-3
    om = O.OMap(pos, dirs, wl, power)
    q = np.array([[0.5, 0.0, 0.5], [0.41, 0.0, 0.63]])
    nrm = np.array([[0.0, 1.0, 0.0]] * 2)
    got = om.gather(q, nrm, 0.06 ** 2, 0, f32set=True)
    assert np.allclose(got[:, 2], n_side * n_side * power, rtol=0.02)
    assert np.all(got[:, :2] == 0)
