"""Closed forms that pin the oracle's eye-tree combine, reflection_glossy,
the TS surface's photon branches and the one-sample lights' direct light.
CPU only.

Each check is chosen so that a plausible slip in the restatement fails it:
  - eye tree (tracer.rs:146-176 with Surface::bsdf, surface.rs:135-206):
    a Lambertian node over a uniform photon carpet is rho/pi * sigma * P (the
    1/pi of estimate_radiance and the 1/pi of the BSDF both); every hit adds
    emittance/(2 pi); a dielectric node's children carry F and 1 - F
    (Schlick), so equal emitters on both sides give exactly emittance/(2 pi);
    F is taken at cos0 = min(cos1, cos2), which at 60 degrees on IOR 1.5 is
    cos1, and for a material of IOR 0 (no refraction: cos2 = 0) makes F = 1
    whatever F0;
  - reflection_glossy (physics.rs:232-261): the sample's component along the
    mirror direction is y = xi^(pw c0), so E[y] = 1 / (1 + pw c0) and
    y^(1/(pw c0)) is uniform; every sample lies on the normal's side;
  - Surface::next_direction for TS (surface.rs:211-258): branch frequencies
    f, (1-f)(1-a), (1-f) a s, (1-f) a (1-s) with f Schlick at min(cos1, cos2);
    under total internal reflection f = 1;
  - the direct light of a sun or point light (light.rs:93-150,
    tracer.rs:263-290): get_radiance puts RADIANCE0 first and the zip with
    the cosines pairs it with the light's only sample, so the sum is exactly
    zero even where the point is lit; the sun's get_radiance (colour * flux,
    light.rs:145-146) therefore never reaches an image.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from ppmpa_amd import scene as S

SR_HALF = 1.0 / (2.0 * math.pi)


def _scene(materials, shapes, lights, camera):
    sc = S.Scene()
    sc.materials, sc.shapes, sc.lights, sc.camera = materials, shapes, lights, camera
    sc.bounds_min, sc.bounds_max = (-2.0, -2.0, -2.0), (2.0, 2.0, 2.0)
    sc.bvh = 1  # every shape tested one by one, as the reference does
    return sc


def _far_map():
    """A map whose one photon is far from every query: di = 0."""
    return O.OMap([[50.0, 50.0, 50.0]], [[0.0, -1.0, 0.0]], [0], 1.0)


def _eye(sc, omap=None, r2=1e-4, maxd=6, uc=False, seed=5):
    return O.OScene(sc).trace_eye(omap or _far_map(), seed, 1, r2, uc, maxd, 0)


def _down_camera(height=0.5, fov=1.0, tilt_deg=0.0, n=8):
    """Camera over (0.5, 0, 0.5) looking at it, tilt_deg from the floor normal."""
    t = math.radians(tilt_deg)
    eye = (0.5, height * math.cos(t), 0.5 - height * math.sin(t))
    up = (0.0, 0.0, 1.0) if tilt_deg == 0.0 else (0.0, 1.0, 0.0)
    return S.pinhole(n, n, eye, (0.5, 0.0, 0.5), fov_deg=fov, upper=up)


def _emitter(e):
    """A black Lambertian emitter: its node adds emittance/(2 pi) and nothing else."""
    return S.simple((0, 0, 0), (0, 0, 0), 1.0, 0.0, 0.0, emittance=e)


FLOOR = S.plain((0, 1, 0), 0.0, 0)


def test_lambertian_over_uniform_carpet():
    """rho/pi * (P * count / (pi r^2)) ~= rho/pi * sigma_c * P per channel,
    plus the floor's own emittance/(2 pi)."""
    rho = (0.6, 0.3, 0.9)
    em = (0.2, 0.0, 0.5)
    mats = [S.simple(rho, (0, 0, 0), 1.0, 0.0, 0.0, emittance=em)]
    sc = _scene(mats, [FLOOR], [], _down_camera(0.4, fov=20.0, n=16))
    n_side = 600
    xs = (np.arange(n_side) + 0.5) / n_side
    X, Z = np.meshgrid(xs, xs)
    pos = np.stack([X.ravel(), np.zeros(X.size), Z.ravel()], 1)
    ij = np.add.outer(np.arange(n_side), np.arange(n_side)).ravel()
    wl = (ij % 3).astype(np.uint8)
    power = 1e-4
    om = O.OMap(pos, np.tile([0.0, -1.0, 0.0], (len(wl), 1)), wl, power)
    img = _eye(sc, om, r2=0.03 ** 2)
    sigma_c = n_side * n_side / 3.0  # photons per m^2 per channel
    want = np.array(rho) / math.pi * sigma_c * power + np.array(em) * SR_HALF
    got = img.reshape(-1, 3).mean(axis=0)
    assert np.allclose(got, want, rtol=0.01), (got, want)
    # per pixel within the lattice count's spread
    assert np.allclose(img.reshape(-1, 3), want, rtol=0.06)


def test_emitter_seen_directly():
    em = (3.0, 2.0, 1.0)
    sc = _scene([_emitter(em)], [FLOOR], [], _down_camera())
    img = _eye(sc)
    assert np.allclose(img, np.array(em) * SR_HALF, rtol=1e-12, atol=0)


def _two_emitters(floor_mat, e_up, e_down, tilt=0.0, fov=1.0):
    """floor_mat on y = 0, an emitter on y = 1 facing down (e_up) and one on
    y = -1 facing up (e_down); the camera at height 0.5 looks at the floor."""
    mats = [floor_mat, _emitter(e_up), _emitter(e_down)]
    shapes = [FLOOR, S.plain((0, -1, 0), 1.0, 1), S.plain((0, 1, 0), 1.0, 2)]
    return _scene(mats, shapes, [], _down_camera(0.5, fov=fov, tilt_deg=tilt))


def schlick(f0, c):
    return f0 + (1.0 - f0) * (1.0 - c) ** 5


def test_metal_of_ior_zero_reflects_everything():
    """IOR 0: specular_refraction finds no direction and cos2 = 0, so
    cos0 = min(cos1, 0) = 0 and F = schlick(F0, 0) = 1 whatever F0."""
    e = np.array([1.0, 2.0, 4.0])
    metal = S.simple((0, 0, 0), (0.3, 0.3, 0.3), 0.0, 1.0, 0.0)
    img = _eye(_two_emitters(metal, e, (0, 0, 0)))
    assert np.allclose(img, e * SR_HALF, rtol=1e-12)


@pytest.mark.parametrize("tilt", [0.0, 60.0])
def test_metal_of_ior_1_5_takes_f_at_min_cos(tilt):
    """IOR 1.5: cos2 = sqrt(2.25 + cos1^2 - 1) / 1.5 >= cos1, so F is Schlick
    at cos1 (0.3 at normal incidence; 0.3 + 0.7 / 32 at 60 degrees)."""
    e = np.array([1.0, 2.0, 4.0])
    metal = S.simple((0, 0, 0), (0.3, 0.3, 0.3), 0.0, 1.0, 0.0, ior=(1.5, 1.5, 1.5))
    img = _eye(_two_emitters(metal, e, (5.0, 5.0, 5.0), tilt=tilt, fov=0.05))
    c1 = math.cos(math.radians(tilt))
    c2 = math.sqrt(1.5 ** 2 + c1 ** 2 - 1.0) / 1.5
    assert c1 <= c2
    want = schlick(0.3, min(c1, c2)) * e * SR_HALF  # metal: no transmitted term
    assert np.allclose(img, want, rtol=1e-3), (img.reshape(-1, 3)[0], want)
    # the larger cosine would give a weight visibly off at 60 degrees
    if tilt:
        assert abs(schlick(0.3, c2) - schlick(0.3, c1)) > 0.02


def test_dielectric_children_sum_to_one():
    """Equal emitters on both sides of a glass plane: F E + (1 - F) E, plus
    the glass's own emittance, whatever the angle."""
    e = np.array([1.5, 0.5, 2.5])
    eg = np.array([0.1, 0.2, 0.3])
    glass = S.simple((0, 0, 0), (0.04, 0.04, 0.04), 0.0, 0.0, 0.0, ior=(1.5, 1.5, 1.5), emittance=eg)
    for tilt in (0.0, 35.0, 60.0):
        img = _eye(_two_emitters(glass, e, e, tilt=tilt, fov=10.0))
        assert np.allclose(img, (e + eg) * SR_HALF, rtol=1e-9), tilt


@pytest.mark.parametrize("tilt", [0.0, 60.0])
def test_dielectric_reflection_and_transmission_weights(tilt):
    """Glass (F0 = 0.04, IOR 1.5): the reflected child carries F, the
    transmitted one 1 - F, F at min(cos1, cos2) = cos1 from air."""
    e = np.array([1.0, 2.0, 4.0])
    glass = S.simple((0, 0, 0), (0.04, 0.04, 0.04), 0.0, 0.0, 0.0, ior=(1.5, 1.5, 1.5))
    c1 = math.cos(math.radians(tilt))
    f = schlick(0.04, c1)
    up = _eye(_two_emitters(glass, e, (0, 0, 0), tilt=tilt, fov=0.005))
    down = _eye(_two_emitters(glass, (0, 0, 0), e, tilt=tilt, fov=0.005))
    assert np.allclose(up, f * e * SR_HALF, rtol=1e-3)
    assert np.allclose(down, (1.0 - f) * e * SR_HALF, rtol=1e-4)


def test_diffuse_weight_excludes_specular_children():
    """A half-diffuse node (diffuseness 0.5, dielectric, IOR 0): 0.5 rho/pi di
    + 0.5 (F si + (1 - F) ti) with F = 1 (IOR 0), so the transmitted side's
    emitter never shows."""
    rho = (0.8, 0.8, 0.8)
    mat = S.simple(rho, (0.1, 0.1, 0.1), 0.5, 0.0, 0.0)
    e_up = np.array([2.0, 1.0, 0.5])
    img = _eye(_two_emitters(mat, e_up, (9.0, 9.0, 9.0)))
    assert np.allclose(img, 0.5 * e_up * SR_HALF, rtol=1e-12)


# ---------------------------------------------------------------- glossy

@pytest.mark.parametrize("pw", [0.5, 3.0, 20.0])
def test_reflection_glossy_moments(pw):
    from scipy import stats
    n = np.array([0.0, 1.0, 0.0])
    r = np.array([0.5, 1.0, 0.3])
    r /= np.linalg.norm(r)
    c0 = float(n @ r)
    a = pw * c0
    N = 100000
    w = O.reflection_glossy(n, r, pw, seed=9, count=N)
    assert np.allclose(np.linalg.norm(w, axis=1), 1.0, atol=1e-12)
    assert np.all(w @ n >= 0.0)
    y = w @ r
    mean, var = 1.0 / (1.0 + a), 1.0 / (2.0 * a + 1.0) - 1.0 / (1.0 + a) ** 2
    assert abs(y.mean() - mean) < 5.0 * math.sqrt(var / N), (y.mean(), mean)
    # y = xi^a is uniform in xi: y^(1/a) ~ U(0, 1).  Below y ~ 1e-16 the dot
    # product cannot recover y, so test the range y > 1e-12, u in [u0, 1].
    u0 = 1e-12 ** (1.0 / a)
    sel = y > 1e-12
    u = y[sel] ** (1.0 / a)
    assert abs(sel.mean() - (1.0 - u0)) < 5.0 * math.sqrt(u0 * (1.0 - u0) / N) + 1e-5
    assert stats.kstest(u, "uniform", args=(u0, 1.0 - u0)).pvalue > 1e-4


def test_reflection_glossy_mirror_limit():
    """roughness 0: pw = 1/(10^5 + 1), the samples stay within a few
    milliradians of the mirror direction."""
    n = np.array([0.0, 1.0, 0.0])
    r = np.array([0.0, 0.6, 0.8])
    w = O.reflection_glossy(n, r, 1.0 / (1e5 + 1.0), seed=3, count=20000)
    ang = np.arccos(np.clip(w @ r, -1, 1))
    assert ang.max() < 0.02 and np.median(ang) < 4e-3


# ---------------------------------------------------------------- TS branches

def _classify(kinds, dirs, mirror):
    spec = (kinds == 1) & (dirs @ mirror > 0.999)
    return dict(ref=spec, dif=(kinds == 1) & ~spec, tr=kinds == 2, abs=kinds == 0)


@pytest.mark.parametrize("wl", [0, 1, 2])
def test_ts_branch_frequencies(wl):
    a_d = (0.9, 0.6, 0.3)
    a_s = (0.1, 0.3, 0.5)
    scat = 0.4
    m = S.ts(a_d, a_s, scat, 0.0, 0.0, ior=(1.5, 1.5, 1.5))
    n = np.array([0.0, 1.0, 0.0])
    v = np.array([0.3, -1.0, 0.2])
    v /= np.linalg.norm(v)
    N = 200000
    kinds, dirs = O.next_direction(m, 1.5, n, v, wl, seed=17, count=N)
    c1 = -float(v @ n)
    c2 = math.sqrt(1.5 ** 2 + c1 ** 2 - 1.0) / 1.5
    f = schlick(a_s[wl], min(c1, c2))
    p = dict(ref=f, abs=(1 - f) * (1 - a_d[wl]), dif=(1 - f) * a_d[wl] * scat, tr=(1 - f) * a_d[wl] * (1 - scat))
    mirror = v + 2.0 * c1 * n
    got = _classify(kinds, dirs, mirror)
    for k, pk in p.items():
        frac = got[k].mean()
        assert abs(frac - pk) < 5.0 * math.sqrt(pk * (1 - pk) / N) + 1e-4, (k, frac, pk)
    # diffuse: uniform over the hemisphere (E[cos] = 1/2, E[cos^2] = 1/3)
    cd = dirs[got["dif"]] @ n
    assert np.all(cd > 0.0)
    assert abs(cd.mean() - 0.5) < 0.01 and abs((cd ** 2).mean() - 1.0 / 3.0) < 0.01
    # transmitted: Snell about the normal, t = (v + (c1 - g) n) / eta
    g = math.sqrt(1.5 ** 2 + c1 ** 2 - 1.0)
    t = (v + (c1 - g) * n) / 1.5
    t /= np.linalg.norm(t)
    assert np.all(dirs[got["tr"]] @ t > 0.9999)


def test_ts_total_internal_reflection_always_reflects():
    """From glass into air beyond the critical angle there is no refracted
    direction: cos2 = 0, so F = schlick(F0, 0) = 1 and every draw reflects."""
    m = S.ts((0.9, 0.9, 0.9), (0.04, 0.04, 0.04), 0.3, 0.0, 0.0, ior=(1.0, 1.0, 1.0))
    n = np.array([0.0, 1.0, 0.0])
    v = np.array([0.95, -0.3, 0.0])
    v /= np.linalg.norm(v)
    kinds, dirs = O.next_direction(m, 1.0 / 1.5, n, v, 1, seed=4, count=5000)
    assert np.all(kinds == 1)
    mirror = v + 2.0 * (-(v @ n)) * n
    assert np.all(dirs @ mirror > 0.999)


# ---------------------------------------------------------------- one-sample lights

def _lit_floor_scene(light):
    mats = [S.lambertian(0.7), S.lambertian(0.0)]
    panel = S.parallelogram((-1.0, 0.9, -1.0), (2.0, 0.9, -1.0), (-1.0, 0.9, 2.0), 1)
    return _scene(mats, [FLOOR, panel], [light], _down_camera(0.4))


def test_sun_and_point_light_direct_light_is_zero():
    """One sample per light, paired with RADIANCE0 by the zip: exactly zero,
    although the floor point is lit (a parallelogram light in the same place
    gives light there)."""
    p, n = (0.5, 0.0, 0.5), (0.0, 1.0, 0.0)
    d = np.array([1.0, -2.0, 0.5])
    d /= np.linalg.norm(d)
    sun = S.sun_light((1, 1, 1), 50.0, (-1.0, 0.9, -1.0), (0, -1, 0), (3.0, 0, 0), (0, 0, 3.0), d)
    point = S.point_light((1, 1, 1), 50.0, (0.5, 0.8, 0.5))
    para = S.parallelogram_light((1, 1, 1), 50.0, (0.3, 0.85, 0.3), (0, -1, 0), (0.4, 0, 0), (0, 0, 0.4))
    for light in (sun, point):
        assert np.all(O.OScene(_lit_floor_scene(light)).direct_light(p, n) == 0.0)
    assert np.all(O.OScene(_lit_floor_scene(para)).direct_light(p, n) > 0.0)
