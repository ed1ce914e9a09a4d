"""Closed forms that pin the oracle's trace_ray_classic (tracer.rs:221-259),
the tree of rtc (rtc.rs:17-47).  CPU only.

Each check is chosen so that a plausible slip in the restatement fails it:
  - di = the direct light plus the camera's ambient: a Lambertian of
    reflectance rho under ambient A alone shows rho A / pi (the BSDF's 1/pi),
    and a point light adds exactly nothing (RADIANCE0 first, light.rs:131-147);
  - every hit adds emittance / (2 pi);
  - the depth limit is 10 levels (tracer.rs:222): between two parallel
    mirrors of diffuseness d the pixel is rho A / pi * (1 - (1 - d)^10);
  - the BSDF at cos1 and the refraction about the surface normal
    (tracer.rs:232,246,257): through a glass slab at 60 degrees onto an
    emitter, E / (2 pi) (1 - F1)(1 - F2) (1 + F2^2 + F2^4 + F2^6) with F1 =
    Schlick at cos 60 and F2 at the Snell angle inside (the slab's faces are
    both glass, so the ray leaves the lower face unbent, eta = 1 there), four
    internal round trips fitting in the 10 levels.
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
    sc.bvh = 1
    return sc


def _down_camera(height=0.5, fov=1.0, tilt_deg=0.0, n=8):
    t = math.radians(tilt_deg)
    eye = (0.5, height * math.cos(t), 0.5 - height * math.sin(t))
    up = (0.0, 0.0, 1.0) if tilt_deg == 0.0 else (0.0, 1.0, 0.0)
    return S.pinhole(n, n, eye, (0.5, 0.0, 0.5), fov_deg=fov, upper=up)


def _emitter(e):
    return S.simple((0, 0, 0), (0, 0, 0), 1.0, 0.0, 0.0, emittance=e)


def _classic(sc, ambient, seed=3):
    return O.OScene(sc).trace_classic(seed, 1, ambient)


def schlick(f0, c):
    return f0 + (1.0 - f0) * (1.0 - c) ** 5


@pytest.mark.parametrize("with_point", [False, True])
def test_ambient_on_lambertian(with_point):
    rho = np.array([0.6, 0.3, 0.9])
    amb = np.array([0.2, 0.5, 1.0])
    lights = [S.point_light((1, 1, 1), 50.0, (0.5, 0.3, 0.5))] if with_point else []
    sc = _scene([S.simple(rho, (0, 0, 0), 1.0, 0.0, 0.0)], [S.plain((0, 1, 0), 0.0, 0)], lights, _down_camera())
    img = _classic(sc, amb)
    assert np.allclose(img, rho * amb / math.pi, rtol=1e-12, atol=0)


def test_emitter_seen_directly():
    em = np.array([3.0, 2.0, 1.0])
    sc = _scene([_emitter(em)], [S.plain((0, 1, 0), 0.0, 0)], [], _down_camera())
    assert np.allclose(_classic(sc, (0.0, 0.0, 0.0)), em * SR_HALF, rtol=1e-12, atol=0)


@pytest.mark.parametrize("d", [0.3, 0.05])
def test_mirror_pair_depth_limit(d):
    rho = np.array([0.8, 0.5, 0.2])
    amb = np.array([1.0, 1.0, 1.0])
    m = S.simple(rho, (1, 1, 1), d, 1.0, 0.0)  # c2 white: F = 1 at any angle
    sc = _scene([m], [S.plain((0, 1, 0), 0.0, 0), S.plain((0, -1, 0), 1.0, 0)], [], _down_camera(fov=0.5))
    img = _classic(sc, amb)
    want = rho * amb / math.pi * (1.0 - (1.0 - d) ** 10)
    assert np.allclose(img, want, rtol=1e-12), (img.reshape(-1, 3)[0], want)
    # one level more or fewer is visible
    for n in (9, 11):
        assert not np.allclose(img.reshape(-1, 3)[0], rho * amb / math.pi * (1.0 - (1.0 - d) ** n), rtol=1e-6)


def test_glass_slab_at_60_degrees():
    f0, ior = 0.04, 1.5
    e = np.array([1.0, 2.0, 4.0])
    glass = S.simple((0, 0, 0), (f0, f0, f0), 0.0, 0.0, 0.0, ior=(ior, ior, ior))
    # the camera stands 0.25 above the emitter at y = 0; the slab fills 0.08 < y < 0.12
    shapes = [S.plain((0, 1, 0), -0.12, 0), S.plain((0, 1, 0), -0.08, 0), S.plain((0, 1, 0), 0.0, 1)]
    sc = _scene([glass, _emitter(e)], shapes, [], _down_camera(0.5, fov=0.05, tilt_deg=60.0))
    img = _classic(sc, (0.0, 0.0, 0.0))
    c1 = math.cos(math.radians(60.0))
    ct = math.sqrt(1.0 - (math.sin(math.radians(60.0)) / ior) ** 2)
    f1, f2 = schlick(f0, c1), schlick(f0, ct)
    want = e * SR_HALF * (1.0 - f1) * (1.0 - f2) * sum(f2 ** (2 * k) for k in range(4))
    assert np.allclose(img, want, rtol=1e-3), (img.reshape(-1, 3)[0], want)
    # the PPM tree's Fresnel at min(cos1, cos2) or a fifth round trip would differ
    wrong = e * SR_HALF * (1.0 - f1) * (1.0 - f2) * sum(f2 ** (2 * k) for k in range(5))
    assert not np.allclose(img.reshape(-1, 3)[0], wrong, rtol=1e-6)


def test_glass_to_air_interface_takes_f_at_cos1_and_medium_by_side():
    """The slab's lower face is a glass-to-air interface (a material of IOR
    1): leaving the glass the ray bends back to 60 degrees and lands on a
    small emitter 0.46 m along z, where only that exit direction arrives.
    The lower face's Fresnel is Schlick at cos1 (the Snell angle inside),
    not at min(cos1, cos2) = cos 60 as in the PPM tree; a child taking air
    as its medium would not bend there (eta 1) and miss the emitter."""
    f0 = 0.04
    e = np.array([1.0, 2.0, 4.0])
    glass = S.simple((0, 0, 0), (f0, f0, f0), 0.0, 0.0, 0.0, ior=(1.5, 1.5, 1.5))
    out = S.simple((0, 0, 0), (f0, f0, f0), 0.0, 0.0, 0.0, ior=(1.0, 1.0, 1.0))
    shapes = [S.plain((0, 1, 0), -0.12, 0), S.plain((0, 1, 0), -0.08, 1),
              S.parallelogram((0.4, 0.0, 0.42), (0.6, 0.0, 0.42), (0.4, 0.0, 0.50), 2)]
    sc = _scene([glass, out, _emitter(e)], shapes, [], _down_camera(0.5, fov=0.05, tilt_deg=60.0))
    img = _classic(sc, (0.0, 0.0, 0.0))
    c1 = math.cos(math.radians(60.0))
    ct = math.sqrt(1.0 - (math.sin(math.radians(60.0)) / 1.5) ** 2)
    want = e * SR_HALF * (1.0 - schlick(f0, c1)) * (1.0 - schlick(f0, ct))
    assert np.allclose(img, want, rtol=1e-3), (img.reshape(-1, 3)[0], want)
    wrong = e * SR_HALF * (1.0 - schlick(f0, c1)) * (1.0 - schlick(f0, c1))
    assert abs(want[0] - wrong[0]) > 1e-2 * want[0]
