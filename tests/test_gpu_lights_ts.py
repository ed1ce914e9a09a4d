"""GPU parity for the parts of the path the benchmark scenes leave out: TS
surfaces (surface.rs:56-66; metallic with roughness, a rough transparent
glass, a dielectric mix of diffuse and specular), point lights and sun
lights (light.rs:30-37), photon filters, and a blurred (thin-lens) camera.
Each is checked against the FP64 oracle on common random numbers: photon
walks record by record, one eye pass pixel by pixel, and a short PPM-PA run.
"""
import numpy as np
import pytest

import oracle_lib as O
from gpu_common import assert_crn, check_eye_crn, check_photon_trace_crn, ctx, say

import ppmpa_amd as P
from ppmpa_amd import _abi as A
from ppmpa_amd import scene as S

pytestmark = pytest.mark.gpu


def ts_room(res=48):
    """The reference-style room: Simple walls, a silver Simple ball, a TS
    metal ball with roughness 0.5, a rough TS glass ball, the ceiling panel
    light; read_camera's thin-lens camera (blur on) at res x res."""
    sc = P.reference_room()
    c = sc.camera
    sc.camera = S.Camera(res, res, c.eye, c.target, c.upper, focus=c.focus, focal_len=c.focal_len,
                         f_number=c.f_number, progressive=True, antialias=True, blur=True)
    return sc


def _unit_room(res):
    sc = S.Scene()
    white = sc.add_material(S.lambertian(0.7))
    red = sc.add_material(S.lambertian((0.7, 0.2, 0.2)))
    sc.shapes += [S.plain((0, 1, 0), 0.0, white), S.plain((0, -1, 0), 1.0, white), S.plain((0, 0, -1), 1.0, white),
                  S.plain((1, 0, 0), 0.0, red), S.plain((-1, 0, 0), 1.0, white)]
    sc.camera = S.pinhole(res, res, (0.5, 0.5, -1.2), (0.5, 0.45, 0.5), fov_deg=50.0)
    sc.bounds_min, sc.bounds_max = (-0.01, -0.01, -1.21), (1.01, 1.01, 1.01)
    return sc


def point_room(res=48):
    """Two coloured point lights over a dielectric TS ball (diffuse base with
    a specular coat, scatterness 0.6) and a Lambertian ball."""
    sc = _unit_room(res)
    coat = sc.add_material(S.ts((0.2, 0.4, 0.8), (0.04, 0.04, 0.04), 0.6, 0.0, 0.2))
    lam = sc.add_material(S.lambertian(0.5))
    sc.shapes += [S.sphere((0.35, 0.2, 0.55), 0.2, coat), S.sphere((0.72, 0.15, 0.35), 0.15, lam)]
    sc.lights += [S.point_light((1.0, 0.8, 0.6), 3.0, (0.3, 0.9, 0.4)), S.point_light((0.4, 0.6, 1.0), 2.0,
                                                                                          (0.75, 0.85, 0.7))]
    return sc


def sun_room(res=48):
    """A sun light under the ceiling: its parallelogram emits every photon
    along one direction, SUN_DIR; a Simple glass ball in the beam casts a
    caustic on the floor."""
    sc = _unit_room(res)
    glass = sc.add_material(S.glass(1.5))
    sc.shapes.append(S.sphere((0.6, 0.2, 0.6), 0.12, glass))
    sc.lights.append(S.sun_light((1, 1, 1), 4.0, (0.25, 0.999, 0.25), (0, -1, 0), (0.4, 0, 0), (0, 0, 0.4),
                                 tuple(SUN_DIR)))
    return sc


SUN_DIR = np.array([0.2, -1.0, 0.2]) / np.sqrt(1.08)


SCENES = {"ts_room": (ts_room, 0.1), "point_room": (point_room, 0.03), "sun_room": (sun_room, 0.03)}


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("uc", [True, False])
def test_photon_trace_crn(name, uc):
    say("lights/ts photon crn", name, uc)
    sc = SCENES[name][0]()
    kinds = {l.kind for l in sc.lights}
    assert kinds
    check_photon_trace_crn(ctx(), sc, 11, 2, 20000, uc, 10)


@pytest.mark.parametrize("name,pfilter", [("ts_room", A.FILTER_NONE), ("point_room", A.FILTER_CONE),
                                          ("sun_room", A.FILTER_GAUSS), ("ts_room", A.FILTER_GAUSS)])
def test_eye_pass_crn(name, pfilter):
    say("lights/ts eye crn", name, pfilter)
    make, r = SCENES[name]
    check_eye_crn(ctx(), make(), 5, 40000, r * r, True, 10, pfilter=pfilter,
                  kind="eye_glossy" if name == "ts_room" else None)


@pytest.mark.parametrize("name", list(SCENES))
def test_session_crn(name):
    say("lights/ts session crn", name)
    make, r = SCENES[name]
    sc = make(40)
    c = ctx()
    ds = P.DeviceScene(c, sc)
    cfg = P.pass_cfg(seed=4, n_photon=20000, radius0=r, alpha=2 / 3, max_depth=10)
    mean, _, t = P.run_passes(c, ds, cfg, 3)
    omean, _, stored, _, _ = O.OScene(sc).run_passes(cfg, 3)
    assert t["passes"] == 3
    assert abs(t["photons_stored"] - stored) <= 0.01 * stored
    assert_crn(mean, omean, "session_glossy" if name == "ts_room" else "session", f"session {name}")


def test_sun_photons_share_one_direction():
    """Every depth-0 photon of a sun light travels along its direction: with
    use_classic off they are stored on their first hit with the incoming
    direction (tracer.rs:73-78), i.e. the light's own."""
    say("sun direction")
    c = ctx()
    sc = sun_room()
    ds = P.DeviceScene(c, sc)
    n = 5000
    ph = P.trace_photons(c, ds, 3, 1, n, use_classic=False, max_depth=1)
    pos, dir, wl, meta = ph.download()
    # the glass ball's shadow (about 28% of the beam) stores nothing at depth 0
    assert 0.55 * n < len(dir) < 0.9 * n
    assert np.abs(dir - SUN_DIR).max() < 1e-5
    assert np.all(np.abs(pos[:, 1]) < 1e-4)  # all on the floor
