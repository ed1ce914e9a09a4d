"""GPU: pass pipelines.  A session with K pipelines runs pass p on pipeline
p mod K with its own buffers and FP64 sums; every pass is computed exactly as
in one pipeline (its radius is the same FP64 product), so images agree with
the single-pipeline run up to the rounding of the merged sums, and splitting
the run() calls changes nothing."""
import numpy as np
import pytest

from gpu_common import ctx, say

import ppmpa_amd as P

pytestmark = pytest.mark.gpu


def _cfg(k):
    return P.pass_cfg(seed=5, n_photon=30000, radius0=0.02, alpha=2 / 3, max_depth=10, pipelines=k)


def _run(c, ds, k, chunks):
    s = P.Session(c, ds, _cfg(k))
    for n in chunks:
        s.run(n)
    m, v, n = s.image()
    r2 = s.radius2()
    t = s.timings()
    s.close()
    return m, v, n, r2, t


def test_pipelines_match_one_pipeline():
    say("pipelines 1 vs 2 vs 3")
    c = ctx()
    ds = P.DeviceScene(c, P.two_spheres_scene(64, 48))
    m1, v1, n1, r1, t1 = _run(c, ds, 1, [7])
    for k in (2, 3):
        m, v, n, r, t = _run(c, ds, k, [7])
        assert n == n1 == 7
        assert r == r1  # the last pass's radius, the same FP64 product
        assert np.allclose(m, m1, rtol=1e-12, atol=1e-300)
        assert np.allclose(v, v1, rtol=1e-6, atol=1e-10 * float((m1 ** 2).max()))
        assert t["photons_stored"] == t1["photons_stored"]
        assert t["gather_queries"] == t1["gather_queries"]
        print("pipelines", k, "max rel diff", float(np.abs(m - m1).max() / np.abs(m1).max()))


def test_split_runs_are_identical():
    say("pipelines split runs")
    c = ctx()
    ds = P.DeviceScene(c, P.caustic_scene(48, 40))
    a = _run(c, ds, 2, [7])
    b = _run(c, ds, 2, [3, 1, 3])
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[3] == b[3]


def test_reset_and_default_pipelines():
    say("pipelines reset")
    c = ctx()
    ds = P.DeviceScene(c, P.two_spheres_scene(32, 32))
    s = P.Session(c, ds, _cfg(3))
    s.run(5)
    m0, _, _ = s.image(with_var=False)
    s.reset()
    s.run(5)
    m1, _, n = s.image(with_var=False)
    s.close()
    assert n == 5 and np.array_equal(m0, m1)
    with pytest.raises(P.PpmError):
        P.Session(c, ds, _cfg(5))


@pytest.mark.parametrize("n_pass,pipelines", [(512, 1), (512, 2), (1000, 3)])
def test_radius_schedule_closed_form(n_pass, pipelines):
    """r_i^2 = r_1^2 prod_{k=1}^{i-1} (k + alpha)/(k + 1) (util/iterator.rb:33-38),
    kept as an FP64 product on the device: after 512 (1000) passes it equals
    the exact rational product rounded once to 1e-12 relative (measured
    8e-15 and 1.8e-14), and the closed form Gamma(i + alpha) /
    (Gamma(1 + alpha) Gamma(i + 1)) to 1e-10 (lgamma's own precision).  For
    comparison a float32 product would have drifted by ~2e-7 (5e-7); the
    kernels' float r2 is the double rounded once (6e-8 at most)."""
    import math
    from fractions import Fraction
    say("radius schedule", n_pass, pipelines)
    c = ctx()
    ds = P.DeviceScene(c, P.two_spheres_scene(16, 16))
    r1 = 0.01
    cfg = P.pass_cfg(seed=1, n_photon=2000, radius0=r1, alpha=2 / 3, max_depth=2, pipelines=pipelines)
    s = P.Session(c, ds, cfg)
    s.run(n_pass)
    r2 = s.radius2()
    s.close()
    exact = Fraction(1, 10000)  # r1^2, alpha = 2/3: (k + 2/3)/(k + 1) = (3k + 2)/(3k + 3)
    for k in range(1, n_pass):
        exact *= Fraction(3 * k + 2, 3 * k + 3)
    assert abs(r2 - float(exact)) <= 1e-12 * float(exact), (r2, float(exact))
    a = 2.0 / 3.0
    gam = r1 * r1 * math.exp(math.lgamma(n_pass + a) - math.lgamma(1 + a) - math.lgamma(n_pass + 1))
    assert abs(r2 - gam) <= 1e-10 * gam
    f = np.float32(r1 * r1)
    for k in range(1, n_pass):
        f = np.float32(f * np.float32((k + a) / (k + 1)))
    print(f"radius^2 after {n_pass} passes: device rel {abs(r2 - float(exact)) / float(exact):.1e}, "
          f"float32 product rel {abs(float(f) - float(exact)) / float(exact):.1e}")
