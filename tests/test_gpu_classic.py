"""GPU: the classic ray tracer of rtc (trace_ray_classic, tracer.rs:221-259;
rtc.rs:17-47) through ppm_trace_classic: the oracle's closed-form scenes
(ambient on a Lambertian, the depth limit between two mirrors, glass slabs
at 60 degrees) to float rounding, the FP64 oracle's image on common camera
streams for rooms with TS materials, glass and a scene BVH, and the rtc
command line."""
import math
import os

import numpy as np
import pytest

import oracle_lib as O
from gpu_common import assert_crn, ctx, say
from test_oracle_classic import SR_HALF, _down_camera, _emitter, _scene, schlick

import ppmpa_amd as P
from ppmpa_amd import cli
from ppmpa_amd import photonmap_io as M

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _gpu(sc, ambient, seed=3):
    c = ctx()
    return P.trace_classic(c, P.DeviceScene(c, sc), seed, 1, ambient)


def test_closed_forms():
    say("classic closed forms")
    rho, amb = np.array([0.6, 0.3, 0.9]), np.array([0.2, 0.5, 1.0])
    sc = _scene([S_lambert(rho)], [P.plain((0, 1, 0), 0.0, 0)], [], _down_camera())
    assert np.allclose(_gpu(sc, amb), rho * amb / math.pi, rtol=1e-6, atol=0)
    # the depth limit: ten levels between two mirrors of diffuseness 0.3
    m = P.simple(rho, (1, 1, 1), 0.3, 1.0, 0.0)
    sc = _scene([m], [P.plain((0, 1, 0), 0.0, 0), P.plain((0, -1, 0), 1.0, 0)], [], _down_camera(fov=0.5))
    assert np.allclose(_gpu(sc, amb), rho * amb / math.pi * (1.0 - 0.7 ** 10), rtol=1e-5)
    # a glass slab at 60 degrees over an emitter: four internal round trips
    f0, e = 0.04, np.array([1.0, 2.0, 4.0])
    glass = P.simple((0, 0, 0), (f0, f0, f0), 0.0, 0.0, 0.0, ior=(1.5, 1.5, 1.5))
    shapes = [P.plain((0, 1, 0), -0.12, 0), P.plain((0, 1, 0), -0.08, 0), P.plain((0, 1, 0), 0.0, 1)]
    sc = _scene([glass, _emitter(e)], shapes, [], _down_camera(0.5, fov=0.05, tilt_deg=60.0))
    c1 = math.cos(math.radians(60.0))
    ct = math.sqrt(1.0 - (math.sin(math.radians(60.0)) / 1.5) ** 2)
    f1, f2 = schlick(f0, c1), schlick(f0, ct)
    want = e * SR_HALF * (1.0 - f1) * (1.0 - f2) * sum(f2 ** (2 * k) for k in range(4))
    assert np.allclose(_gpu(sc, (0, 0, 0)), want, rtol=1e-3)
    # the glass-to-air face: Fresnel at cos1, the exit bent back onto a small emitter
    out = P.simple((0, 0, 0), (f0, f0, f0), 0.0, 0.0, 0.0, ior=(1.0, 1.0, 1.0))
    shapes = [P.plain((0, 1, 0), -0.12, 0), P.plain((0, 1, 0), -0.08, 1),
              P.parallelogram((0.4, 0.0, 0.42), (0.6, 0.0, 0.42), (0.4, 0.0, 0.50), 2)]
    sc = _scene([glass, out, _emitter(e)], shapes, [], _down_camera(0.5, fov=0.05, tilt_deg=60.0))
    assert np.allclose(_gpu(sc, (0, 0, 0)), e * SR_HALF * (1.0 - f1) * (1.0 - f2), rtol=1e-3)


def S_lambert(rho):
    return P.simple(rho, (0, 0, 0), 1.0, 0.0, 0.0)


@pytest.mark.parametrize("name", ["reference_room", "caustic", "bvh_room"])
def test_against_oracle(name):
    say("classic vs oracle", name)
    if name == "reference_room":
        sc, amb = P.reference_room(), (0.01, 0.02, 0.03)
    elif name == "caustic":
        sc, amb = P.caustic_scene(96, 72), (0.0, 0.0, 0.0)
    else:
        rooms = P.batch_scenes(n=8, seed=1, max_side=64, min_side=48)
        sc = [r for r, _, _ in rooms if len(r.shapes) - 6 >= 64][0]
        amb = (0.05, 0.05, 0.05)
    c = ctx()
    ds = P.DeviceScene(c, sc)
    if name == "bvh_room":
        assert ds.stats()[1] > 0
    got = P.trace_classic(c, ds, 11, 2, amb)
    want = O.OScene(sc).trace_classic(11, 2, amb)
    assert want.max() > 0
    # float rays against double ones, on the same camera streams: the rooms
    # of mirror and glass spheres chain up to ten specular bounces, and a ray
    # that meets a sphere edge differently takes another path from there
    assert_crn(got, want, "classic_specular" if name == "bvh_room" else "classic", f"classic {name}")


def test_rtc_cli(capsys, tmp_path):
    say("cli rtc")
    capsys.readouterr()
    scene, scr = os.path.join(GOLD, "box.scene"), os.path.join(GOLD, "box.scr")
    assert cli.rtc_main([scr, scene, "--seed", "4"]) == 0
    img, _ = M.parse_ppmf(capsys.readouterr().out)
    sc, cam, cfg = cli._camera_scene(scene, scr)
    c = ctx()
    want = P.trace_classic(c, P.DeviceScene(c, sc), 4, 1, cfg["ambient"])
    assert img.shape == want.shape and np.allclose(img, want, rtol=1e-6, atol=0) and img.max() > 0
    # RGB lines for a camera that is not progressive
    rgb_scr = tmp_path / "rgb.scr"
    rgb_scr.write_text(open(scr).read().replace("progressive   : yes", "progressive   : no"))
    assert cli.rtc_main([str(rgb_scr), scene]) == 0
    out = capsys.readouterr().out.split("\n")
    assert out[0] == "P3" and out[4] == "255"
