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
