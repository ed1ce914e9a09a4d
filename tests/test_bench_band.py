"""The 3-sigma agreement rule the benchmark reports (bench.band_fraction):
with variances estimated from K passes per side, the fraction of pixel
channels within 3 standard errors follows Student's t with Welch-
Satterthwaite degrees of freedom.  CPU only, synthetic Gaussian passes."""
import numpy as np
import pytest

import bench


@pytest.mark.parametrize("k", [8, 32])
def test_fraction_matches_student_t(k):
    rng = np.random.default_rng(k)
    a = rng.normal(0.0, 1.0, (k, 120, 120, 3))
    b = rng.normal(0.0, 2.0, (k, 120, 120, 3))  # unequal spreads: Welch's df
    frac, exp, sig, n = bench.band_fraction(a.mean(0), a.var(0, ddof=1), b.mean(0), b.var(0, ddof=1), k, k)
    assert n == a[0].size
    assert abs(frac - exp) <= 4 * sig, (frac, exp, sig)
    # far from the normal 0.9973 at small K
    if k == 8:
        assert exp < 0.992


def test_biased_side_is_detected():
    """A mean shift of one per-pass sigma at K = 32 drops the fraction far
    below the expectation."""
    k = 32
    rng = np.random.default_rng(3)
    a = rng.normal(0.0, 1.0, (k, 80, 80, 3))
    b = rng.normal(1.0, 1.0, (k, 80, 80, 3))
    frac, exp, sig, _ = bench.band_fraction(a.mean(0), a.var(0, ddof=1), b.mean(0), b.var(0, ddof=1), k, k)
    assert exp - frac > 20 * sig
