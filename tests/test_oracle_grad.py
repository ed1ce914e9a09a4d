"""The oracle's gradient of the PPM-PA estimator with respect to the
Lambertian albedos (c1), against central finite differences of the oracle's
own loss on the same random numbers.  CPU only.

A diffuse bounce survives with probability q and multiplies the photon's
weight by c1/q (with q = c1 this is the reference's rule, tracer.rs:83-92,
and every weight is 1).  Holding q at the base albedos, the image is a smooth
function of c1 for fixed random numbers, so the FP64 central difference
(h = 1e-4, truncation error O(h^2)) must equal the analytic gradient to
about 1e-6."""
import numpy as np

import oracle_lib as O

import ppmpa_amd as P


def _setup():
    alb = P.inverse_albedos(3)
    sc = P.inverse_scene(alb, 24, 24, seed=3)
    cfg = P.pass_cfg(seed=5, n_photon=3000, radius0=0.03, alpha=2 / 3, max_depth=6)
    return sc, alb, cfg


def _loss(osc, cfg, n, alb, q, target):
    osc.set_params(c1=_full(osc, alb), q=_full(osc, q))
    mean, _, _, _, _ = osc.run_passes(cfg, n)
    return float(((mean - target) ** 2).sum()), mean


def _full(osc, alb32):
    """c1 of all the scene's materials with the first 32 replaced."""
    m = np.array([mt.color1 for mt in osc.scene.materials], np.float64)
    m[:32] = alb32
    return m


def test_oracle_gradient_matches_central_differences():
    sc, alb, cfg = _setup()
    osc = O.OScene(sc)
    n = 2
    target = np.zeros((24, 24, 3)) + 0.5
    L0, mean = _loss(osc, cfg, n, alb, alb, target)
    osc.set_params(c1=_full(osc, alb), q=_full(osc, alb))
    grad, mean2 = osc.run_passes_grad(cfg, n, 2.0 * (mean - target))
    assert np.allclose(mean, mean2, rtol=1e-12, atol=0)
    h = 1e-4
    checked = 0
    for (m, c) in [(0, 0), (0, 2), (1, 0), (2, 1), (3, 0), (7, 2), (15, 1), (31, 0)]:
        up, dn = alb.copy(), alb.copy()
        up[m, c] += h
        dn[m, c] -= h
        _, Ip = _loss(osc, cfg, n, up, alb, target)
        _, Im = _loss(osc, cfg, n, dn, alb, target)
        # L+ - L- summed per pixel as (I+ - I-)(I+ + I- - 2T): no cancellation
        # between two large loss totals
        fd = float(((Ip - Im) * (Ip + Im - 2 * target)).sum()) / (2 * h)
        if abs(fd) < 1e-9 and abs(grad[m, c]) < 1e-9:
            continue
        assert abs(fd - grad[m, c]) <= 1e-6 * max(abs(fd), 1e-3), (m, c, fd, grad[m, c])
        checked += 1
    assert checked >= 6


def test_survival_reweighting_keeps_the_expectation():
    """q != c1 changes the random decisions but not the expected image: the
    same passes with q = 0.5 everywhere agree with q = c1 within noise, and
    the image is unchanged when q = c1 is set explicitly."""
    sc, alb, cfg = _setup()
    osc = O.OScene(sc)
    base, _, _, _, _ = osc.run_passes(cfg, 6)
    osc.set_params(c1=_full(osc, alb), q=_full(osc, alb))
    same, _, _, _, _ = osc.run_passes(cfg, 6)
    assert np.array_equal(base, same)
    osc.set_params(c1=_full(osc, alb), q=np.full((len(sc.materials), 3), 0.5))
    rew, _, _, _, _ = osc.run_passes(cfg, 6)
    a, b = base.mean(), rew.mean()
    assert abs(a - b) < 0.05 * a, (a, b)
