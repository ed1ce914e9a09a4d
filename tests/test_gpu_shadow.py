"""GPU parity of the direct light (tracer.rs:136-144, light.rs:164-170) where
shadow rays meet many shapes: a box lit by its ceiling panel with 40 small
spheres and 6 tilted triangles floating between the floor and the light, so
that many shapes lie near the edges of the shadow pyramids of the floor and
wall points.  The shading kernel drops, per query, the candidate occluders
that no shadow ray can meet before its light sample; a wrong drop shows as
lit pixels inside shadows.  One eye pass over 50 photons (the image is
the direct term) is checked pixel by pixel against the FP64 oracle on common
random numbers, with quad walls and with infinite plane walls.
"""
import numpy as np
import pytest

from gpu_common import check_eye_crn, ctx, say

import ppmpa_amd as P
from ppmpa_amd import scene as S


def clutter_box(res=96, walls="quad", seed=5):
    sc = P.cornell_box(S.Scene(), walls=walls)
    d = sc.add_material(S.lambertian(0.7))
    rng = np.random.default_rng(seed)
    for _ in range(40):
        c = (rng.uniform(0.2, 0.8), rng.uniform(0.3, 0.9), rng.uniform(0.2, 0.8))
        sc.shapes.append(S.sphere(c, rng.uniform(0.02, 0.05), d))
    for _ in range(6):
        o = np.array([rng.uniform(0.2, 0.7), rng.uniform(0.35, 0.85), rng.uniform(0.2, 0.7)])
        a = o + rng.normal(0.0, 0.08, 3)
        b = o + rng.normal(0.0, 0.08, 3)
        sc.shapes.append(S.polygon(tuple(o), tuple(a), tuple(b), d))
    sc.camera = P.two_spheres_scene(res, res).camera
    return sc


@pytest.mark.gpu
@pytest.mark.parametrize("walls", ["quad", "plane"])
def test_direct_light_clutter_crn(walls):
    say("shadow clutter eye crn", walls)
    sc = clutter_box(96, walls)
    got, want = check_eye_crn(ctx(), sc, 7, 50, 0.01 ** 2, True, 6, frac_min=0.99)
    # most pixels see lit or partly shadowed surfaces
    assert (want.max(axis=2) > 0.02).mean() > 0.5


@pytest.mark.gpu
def test_direct_light_reference_room_crn():
    say("shadow reference room eye crn")
    sc = P.reference_room()
    c = sc.camera
    sc.camera = S.Camera(64, 64, c.eye, c.target, c.upper, focus=c.focus, focal_len=c.focal_len,
                         f_number=c.f_number, progressive=True, antialias=True, blur=False)
    check_eye_crn(ctx(), sc, 9, 50, 0.01 ** 2, True, 6, frac_min=0.99)
