"""Pin the oracle's intersection, ball query, gather, direct light and
photon walk with closed forms, brute force and exact invariants.  CPU only."""
import math

import numpy as np
import pytest

import oracle_lib as O
import ppmpa_amd as P
from ppmpa_amd import scene as S


def one_object_scene(shape_fn, cam=None):
    sc = S.Scene()
    m = sc.add_material(S.lambertian(0.5))
    sc.shapes.append(shape_fn(m))
    sc.lights.append(S.point_light((1, 1, 1), 1.0, (0, 5, 0)))
    sc.camera = cam or S.box_camera(8, 8)
    sc.bounds_min, sc.bounds_max = (-10, -10, -10), (10, 10, 10)
    return O.OScene(sc)


def test_plane_intersection_closed_form():
    osc = one_object_scene(lambda m: S.plain((0, 1, 0), -0.25, m))  # y = 0.25
    obj, pos, nrm = osc.intersect((0.3, 1.0, 0.2), (0.0, -1.0, 0.0))
    assert obj == 0 and np.allclose(pos, [0.3, 0.25, 0.2]) and list(nrm) == [0, 1, 0]
    # from below the normal flips to face the ray (tracer.rs:340-342)
    obj, pos, nrm = osc.intersect((0.3, -1.0, 0.2), (0.0, 1.0, 0.0))
    assert obj == 0 and list(nrm) == [0, -1, 0]
    # parallel ray: no hit; hit behind the origin: none
    assert osc.intersect((0, 1, 0), (1, 0, 0))[0] == -1
    assert osc.intersect((0, 1, 0), (0, 1, 0))[0] == -1


def test_sphere_intersection_closed_form():
    osc = one_object_scene(lambda m: S.sphere((0, 0, 0), 0.5, m))
    obj, pos, nrm = osc.intersect((0, 0, -2), (0, 0, 1))
    assert obj == 0 and np.allclose(pos, [0, 0, -0.5]) and np.allclose(nrm, [0, 0, -1])
    # from inside: the far root, normal flipped inward
    obj, pos, nrm = osc.intersect((0, 0, 0), (1, 0, 0))
    assert obj == 0 and np.allclose(pos, [0.5, 0, 0]) and np.allclose(nrm, [-1, 0, 0])
    # grazing miss
    assert osc.intersect((0, 0.6, -2), (0, 0, 1))[0] == -1


@pytest.mark.parametrize("kind", ["polygon", "parallelogram"])
def test_moller_trumbore_inside_outside(kind):
    fn = S.polygon if kind == "polygon" else S.parallelogram
    osc = one_object_scene(lambda m: fn((0, 0, 0), (1, 0, 0), (0, 0, 1), m))
    # (u, v) = (0.25, 0.25): inside both
    obj, pos, _ = osc.intersect((0.25, 1, 0.25), (0, -1, 0))
    assert obj == 0 and np.allclose(pos, [0.25, 0, 0.25])
    # (0.75, 0.75): outside the triangle (u + v > 1), inside the parallelogram (<= 2)
    obj, _, _ = osc.intersect((0.75, 1, 0.75), (0, -1, 0))
    assert obj == (0 if kind == "parallelogram" else -1)
    assert osc.intersect((1.5, 1, 0.5), (0, -1, 0))[0] == -1


def test_nearest_hit_and_epsilon():
    """Nearest t >= NEARLY0 over all objects (tracer.rs:320-336)."""
    sc = S.Scene()
    m = sc.add_material(S.lambertian(0.5))
    sc.shapes += [S.plain((0, 1, 0), -0.5, m), S.plain((0, 1, 0), -0.3, m), S.sphere((0, 2, 0), 0.1, m)]
    sc.lights.append(S.point_light((1, 1, 1), 1.0, (0, 5, 0)))
    sc.camera = S.box_camera(8, 8)
    osc = O.OScene(sc)
    assert osc.intersect((0, 1, 0), (0, -1, 0))[0] == 0  # y=0.5 hit first from above
    # a ray starting on the plane y = 0.5 ignores it (t < 1e-4) and hits y = 0.3
    assert osc.intersect((0, 0.5, 0), (0, -1, 0))[0] == 1
    assert osc.intersect((0, 0.5 + 0.5e-4, 0), (0, -1, 0))[0] == 1
    assert osc.intersect((0, 0.5 + 2e-4, 0), (0, -1, 0))[0] == 0


# ---------------------------------------------------------------------------
# KdTree::within restated: oracle kd-tree == brute force; kdtreetest fixture

def test_kdtreetest_fixture():
    pos = np.array([[0, 0, 0], [1, 1, 0], [1, 1, 1], [-1, -1, -1], [2, 2, 2], [2, 2, 2]], float)
    dir = np.tile([0.0, 1.0, 0.0], (6, 1))
    m = O.OMap(pos, dir, np.zeros(6, np.uint8), 1.0)
    idx, d2 = m.within([0, 0, 0], 2.5)
    assert list(idx) == [0, 1] and list(d2) == [0.0, 2.0]
    idx, d2 = m.within([0, 0, 0], 3.5 * 3.5)
    assert list(idx) == [0, 1, 2, 3, 4, 5] and list(d2) == [0, 2, 3, 3, 12, 12]
    assert [list(x) for x in O.within_brute(pos, [[0, 0, 0]], 2.5)] == [[0, 1]]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_kd_within_matches_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = 3000
    pos = rng.uniform(0, 1, (n, 3))
    pos[:50] = pos[50:100]  # duplicates
    pos[100:150, 0] = 0.5   # many equal coordinates
    dir = np.tile([0.0, -1.0, 0.0], (n, 1))
    m = O.OMap(pos, dir, np.zeros(n, np.uint8), 1.0)
    q = np.concatenate([rng.uniform(-0.1, 1.1, (150, 3)), pos[:50]])
    r2 = 0.08 ** 2
    brute = O.within_brute(pos, q, r2)
    for i in range(len(q)):
        idx, d2 = m.within(q[i], r2)
        assert sorted(idx.tolist()) == brute[i].tolist()
        assert np.all(np.diff(d2) >= 0)  # ascending by distance, as the crate returns


def test_float_predicate_differs_only_at_the_boundary():
    """The float predicate (GPU order, no contraction) may disagree with the
    FP64 one only for photons within float rounding of the sphere."""
    rng = np.random.default_rng(9)
    q = np.array([[0.4, 0.5, 0.6]])
    r = 0.01
    d = rng.normal(size=(5000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pos = q + d * r * (1 + rng.uniform(-1e-6, 1e-6, 5000))[:, None]
    a = set(O.within_brute(pos, q, r * r, 0)[0].tolist())
    b = set(O.within_brute(pos, q, r * r, 1)[0].tolist())
    diff = a ^ b
    assert len(diff) > 0  # the band is populated on purpose
    d2 = ((pos[list(diff)] - q) ** 2).sum(1)
    assert np.all(np.abs(d2 / (r * r) - 1) < 1e-5)


def test_gather_uniform_carpet_closed_form():
    """sigma photons / m^2 at normal incidence: L = Σ P / (pi r^2) ~ sigma P."""
    ns = 300
    xs = (np.arange(ns) + 0.5) / ns
    X, Z = np.meshgrid(xs, xs)
    pos = np.stack([X.ravel(), np.zeros(X.size), Z.ravel()], 1)
    dir = np.tile([0.0, -1.0, 0.0], (pos.shape[0], 1))
    wl = np.full(pos.shape[0], 1, np.uint8)
    P_ = 2e-3
    m = O.OMap(pos, dir, wl, P_)
    got = m.gather([[0.5, 0, 0.5]], [[0, 1, 0]], 0.06 ** 2)
    assert got[0, 0] == 0 and got[0, 2] == 0
    assert abs(got[0, 1] / (ns * ns * P_) - 1) < 0.02
    # photons arriving from below the surface contribute nothing (max(0, -n.dir))
    m2 = O.OMap(pos, -dir, wl, P_)
    assert np.all(m2.gather([[0.5, 0, 0.5]], [[0, 1, 0]], 0.06 ** 2) == 0)


# ---------------------------------------------------------------------------
# direct light: the 25-sample grid and the RADIANCE0 shift (light.rs:93-150,
# tracer.rs:263-290), restated independently here in numpy.

def test_direct_light_off_by_one():
    sc = S.Scene()
    m = sc.add_material(S.lambertian(0.5))
    lm = sc.add_material(S.simple((0, 0, 0), (0.8, 0.8, 0.8), 0.0, 0.0, 0.0, emittance=(1, 1, 1)))
    p0, s = np.array([0.2, 1.0, 0.3]), 0.4
    sc.shapes += [S.plain((0, 1, 0), 0.0, m),
                  S.parallelogram(p0, p0 + [s, 0, 0], p0 + [0, 0, s], lm, nvec=(0, -1, 0))]
    sc.lights.append(S.parallelogram_light((1, 1, 1), 7.0, p0, (0, -1, 0), (s, 0, 0), (0, 0, s)))
    sc.camera = S.box_camera(8, 8)
    osc = O.OScene(sc)
    p, n = np.array([0.45, 0.0, 0.41]), np.array([0.0, 1.0, 0.0])
    got = osc.direct_light(p, n)
    ts = [0.1, 0.3, 0.5, 0.7, 0.9]
    dists, coss = [], []
    for a in ts:
        for b in ts:
            ld = p0 + a * np.array([s, 0, 0]) + b * np.array([0, 0, s]) - p
            cos0 = n @ (ld / np.linalg.norm(ld))
            dists.append(ld @ ld)
            coss.append(cos0 * cos0)
    L0 = [2 * 7.0 * 0.2 * 0.2 / (4 * math.pi * d) for d in dists]
    # radiance list [0, L(d1), ..., L(d25)] zipped with [c1, ..., c25]
    want = sum(l * c for l, c in zip([0.0] + L0, coss)) / 3.0
    assert np.allclose(got, want, rtol=1e-12)
    # without the emissive object nothing is lit (the shadow ray must hit it)
    sc.shapes.pop()
    assert np.all(O.OScene(sc).direct_light(p, n) == 0)


def test_point_light_direct_is_zero():
    """A single sample pairs only with the prepended RADIANCE0."""
    sc = S.Scene()
    m = sc.add_material(S.lambertian(0.5))
    sc.shapes += [S.plain((0, 1, 0), 0.0, m), S.sphere((0, 2, 0), 0.1, m)]
    sc.lights.append(S.point_light((1, 1, 1), 5.0, (0, 2, 0)))
    sc.camera = S.box_camera(8, 8)
    assert np.all(O.OScene(sc).direct_light((0, 0, 0), (0, 1, 0)) == 0)


# ---------------------------------------------------------------------------
# photon walk invariants (tracer.rs:31-125)

def closed_box(albedo):
    sc = S.Scene()
    m = sc.add_material(S.lambertian(albedo))
    sc.shapes += [S.plain((0, 1, 0), 0.0, m), S.plain((0, -1, 0), 1.0, m), S.plain((0, 0, -1), 1.0, m),
                  S.plain((0, 0, 1), 0.0, m), S.plain((1, 0, 0), 0.0, m), S.plain((-1, 0, 0), 1.0, m)]
    sc.lights.append(S.parallelogram_light((1, 1, 1), 10.0, (0.4, 0.999, 0.4), (0, -1, 0), (0.2, 0, 0),
                                           (0, 0, 0.2)))
    sc.camera = S.box_camera(8, 8)
    sc.bounds_min, sc.bounds_max = (0, 0, 0), (1, 1, 1)
    return sc


@pytest.mark.parametrize("maxd", [1, 4, 10])
def test_albedo_one_box(maxd):
    osc = O.OScene(closed_box(1.0))
    pos, dir, wl, meta, pw = osc.trace_photons(1, 1, 2000, True, maxd)
    assert len(wl) == 2000 * (maxd - 1)
    pos, dir, wl, meta, pw = osc.trace_photons(1, 1, 2000, False, maxd)
    assert len(wl) == 2000 * maxd
    assert pw == 10.0 / 2000
    # children are stored before their parent: depth decreases along a path
    pid, depth = meta >> 7, (meta >> 2) & 31
    for k in range(0, 40 * maxd, maxd):
        assert np.all(pid[k:k + maxd] == pid[k])
        assert list(depth[k:k + maxd]) == list(range(maxd - 1, -1, -1))


def test_albedo_zero_box_and_incoming_direction():
    osc = O.OScene(closed_box(0.0))
    assert len(osc.trace_photons(1, 1, 2000, True, 8)[2]) == 0
    pos, dir, wl, meta, _ = O.OScene(closed_box(0.6)).trace_photons(2, 1, 4000, True, 8)
    # a stored direction is the incoming one: it points into the wall hit
    floor = np.abs(pos[:, 1]) < 1e-9
    assert floor.sum() > 100 and np.all(dir[floor, 1] < 0)
    ceil = np.abs(pos[:, 1] - 1) < 1e-9
    assert np.all(dir[ceil, 1] > 0)


def test_channel_frequencies():
    sc = closed_box(1.0)
    sc.lights[0] = S.parallelogram_light((0.2, 0.3, 0.5), 10.0, (0.4, 0.999, 0.4), (0, -1, 0), (0.2, 0, 0),
                                         (0, 0, 0.2))
    _, _, wl, _, _ = O.OScene(sc).trace_photons(3, 1, 30000, True, 2)
    f = np.bincount(wl, minlength=3) / len(wl)
    expect = np.array([0.2, 0.3, 0.5])
    assert np.all(np.abs(f - expect) < 5 * np.sqrt(expect * (1 - expect) / len(wl)))


def test_mirror_photons_never_stored():
    """Simple with diffuseness 0 never stores (surface.rs:298); a mirror floor
    under the light reflects every photon (Schlick F0 = 1)."""
    sc = S.Scene()
    mir = sc.add_material(S.mirror())
    wall = sc.add_material(S.lambertian(0.0))
    sc.shapes += [S.plain((0, 1, 0), 0.0, mir), S.plain((0, -1, 0), 1.0, wall)]
    sc.lights.append(S.parallelogram_light((1, 1, 1), 1.0, (0.4, 0.999, 0.4), (0, -1, 0), (0.2, 0, 0),
                                           (0, 0, 0.2)))
    sc.camera = S.box_camera(8, 8)
    pos, dir, wl, meta, _ = O.OScene(sc).trace_photons(1, 1, 3000, True, 8)
    # the black ceiling (diffuse, albedo 0) stores each photon once at depth 1
    assert len(wl) == 3000 and np.all(np.abs(pos[:, 1] - 1) < 1e-9)
    assert np.all(((meta >> 2) & 31) == 1)


def test_camera_center_ray_is_eye_direction():
    """camera.rs:58-75,244-256: the ray through screen point (lx-1/2,
    ly-1/2) is ez, the direction from the eye to the target."""
    cam = S.Camera(64, 48, (1.0, 2.0, -4.5), (0.0, 1.0, 0.0), blur=False)
    s = cam.screen()
    lx, ly = 64 // 2, 48 // 2
    d = s["origin"] + (lx - 0.5) * s["esx"] + (ly - 0.5) * s["esy"]
    ez = np.array([-1.0, -1.0, 4.5]) / np.linalg.norm([-1.0, -1.0, 4.5])
    assert np.allclose(d / np.linalg.norm(d), ez, atol=1e-15)
    # rows go down: ey = ex x ez has a negative y component for an upright camera
    assert s["esy"][1] < 0
