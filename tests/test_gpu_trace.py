"""GPU parity: photon emission/bouncing and the eye pass against the FP64
oracle on common random numbers (the same Philox streams), plus exact
invariants of the photon walk."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_common import assert_crn, check_photon_trace_crn, ctx, say, small_scene

import ppmpa_amd as P
from ppmpa_amd import scene as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,uc,maxd", [("two_spheres", True, 8), ("caustic", True, 16), ("caustic", False, 10)])
def test_photon_trace_crn(kind, uc, maxd):
    say("photon trace crn", kind, uc, maxd)
    check_photon_trace_crn(ctx(), small_scene(kind), 7, 3, 20000, uc, maxd)


def _closed_box(albedo):
    sc = S.Scene()
    m = sc.add_material(S.lambertian(albedo))
    sc.shapes += [S.plain((0, 1, 0), 0.0, m), S.plain((0, -1, 0), 1.0, m), S.plain((0, 0, -1), 1.0, m),
                  S.plain((0, 0, 1), 0.0, m), S.plain((1, 0, 0), 0.0, m), S.plain((-1, 0, 0), 1.0, m)]
    sc.lights.append(S.parallelogram_light((1, 1, 1), 10.0, (0.4, 0.999, 0.4), (0, -1, 0), (0.2, 0, 0),
                                           (0, 0, 0.2)))
    sc.camera = S.box_camera(16, 16)
    sc.bounds_min, sc.bounds_max = (0, 0, 0), (1, 1, 1)
    return sc


@pytest.mark.parametrize("maxd", [1, 5, 16])
def test_albedo_one_box_stores_depth_minus_one(maxd):
    """Albedo 1, closed box, uc: every photon bounces to the depth cap and is
    stored at depths 1..maxd-1 (tracer.rs:32,77)."""
    say("albedo-1 box", maxd)
    c = ctx()
    ds = P.DeviceScene(c, _closed_box(1.0))
    n = 10000
    ph = P.trace_photons(c, ds, 1, 1, n, use_classic=True, max_depth=maxd)
    assert ph.count() == n * (maxd - 1)
    pos, dir, wl, meta = ph.download()
    depth = (meta >> 2) & 31
    assert depth.min() >= 1 if maxd > 1 else True
    # every stored direction points into the surface it was stored on
    # (incoming direction, tracer.rs:78): inside a closed box the stored
    # position lies on a wall and the direction heads out of the box interior
    if len(pos):
        assert np.all(np.abs(np.linalg.norm(dir, axis=1) - 1) < 1e-5)
    ph_nc = P.trace_photons(c, ds, 1, 1, n, use_classic=False, max_depth=maxd)
    assert ph_nc.count() == n * maxd


def test_albedo_zero_box_stores_nothing():
    say("albedo-0 box")
    c = ctx()
    ds = P.DeviceScene(c, _closed_box(0.0))
    ph = P.trace_photons(c, ds, 1, 1, 10000, use_classic=True, max_depth=8)
    assert ph.count() == 0


def test_channel_frequencies():
    """Channels follow the normalised light colour (light.rs:157-160)."""
    say("channel frequencies")
    c = ctx()
    sc = _closed_box(1.0)
    sc.lights[0] = S.parallelogram_light((0.2, 0.3, 0.5), 10.0, (0.4, 0.999, 0.4), (0, -1, 0), (0.2, 0, 0),
                                         (0, 0, 0.2))
    ds = P.DeviceScene(c, sc)
    n = 60000
    ph = P.trace_photons(c, ds, 3, 1, n, use_classic=True, max_depth=2)
    _, _, wl, _ = ph.download()
    f = np.bincount(wl, minlength=3) / len(wl)
    expect = np.array([0.2, 0.3, 0.5])
    # chi-square-ish 5 sigma bound per channel
    sig = np.sqrt(expect * (1 - expect) / len(wl))
    assert np.all(np.abs(f - expect) < 5 * sig), (f, expect)


def test_eye_pass_crn():
    """One eye pass on the same photon set: GPU float vs oracle FP64."""
    say("eye pass crn")
    sc = small_scene("caustic", 96, 72)
    c = ctx()
    osc = O.OScene(sc)
    pos, dir, wl, meta, pw = osc.trace_photons(5, 1, 40000, True, 16)
    ph = P.PhotonSet(c, len(wl))
    ph.upload(pos, dir, wl)
    ds = P.DeviceScene(c, sc)
    m = P.PhotonMap(c, ph.capacity).build(ph, pw, ds)
    r2 = 0.02 ** 2
    got = P.trace_eye(c, ds, m, seed=9, pass_=1, r2=r2, use_classic=True, max_depth=16)
    om = O.OMap(pos, dir, wl, pw)
    want = osc.trace_eye(om, 9, 1, r2, True, 16)
    assert np.isfinite(got).all()
    assert_crn(got, want, "eye", "eye caustic")


def test_session_matches_oracle_crn():
    """A few PPM-PA passes: device session vs oracle passes (same seed)."""
    say("session crn")
    sc = small_scene("two_spheres", 64, 64)
    c = ctx()
    ds = P.DeviceScene(c, sc)
    cfg = P.pass_cfg(seed=3, n_photon=20000, radius0=0.03, alpha=2 / 3, max_depth=8)
    mean, var, t = P.run_passes(c, ds, cfg, 3)
    osc = O.OScene(sc)
    omean, ovar, stored, _, _ = osc.run_passes(cfg, 3)
    assert t["passes"] == 3
    assert abs(t["photons_stored"] - stored) <= 0.01 * stored
    assert_crn(mean, omean, "session", "session two_spheres")


def test_photon_index_bound():
    """The photon record's index has 25 bits: a pass (or an uploaded set) of
    more than 2^25 photons is refused, not wrapped."""
    from ppmpa_amd import _abi as A
    say("photon index bound")
    c = ctx()
    ds = P.DeviceScene(c, small_scene("two_spheres", 16, 16))
    with pytest.raises(P.PpmError) as e:
        P.trace_photons(c, ds, 1, 1, (1 << 25) + 1, max_depth=1, out=P.PhotonSet(c, 16))
    assert e.value.status == A.PPM_E_RANGE
    with pytest.raises(P.PpmError) as e:
        P.Session(c, ds, P.pass_cfg(seed=1, n_photon=(1 << 25) + 1, radius0=0.01, alpha=2 / 3, max_depth=1))
    assert e.value.status == A.PPM_E_RANGE
