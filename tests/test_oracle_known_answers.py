"""Pin the FP64 oracle: the reference's own asserted values, published
known-answer vectors, and closed forms.  CPU only."""
import math

import numpy as np
import pytest
from scipy.special import gammaln

import oracle_lib as O

L = O.lib()


def d3(*v):
    return np.array(v, dtype=np.float64)


def ptr(a):
    return a.ctypes.data_as(O.dp)


# ---------------------------------------------------------------------------
# tracer.rs:366-377 (test_filter)

def test_filter_cone_reference_values():
    r = 0.1 * 0.1
    assert L.orc_filter_cone(0.0, r) == 2.538461538461538
    assert L.orc_filter_cone(r, r) == 0.23076923076923078


def test_filter_gauss_reference_values():
    """The reference test asserts 1.2730000000000001 and 0.6061526928041553,
    values of the formula with CORR = 0.355 (the constant commented out at
    tracer.rs:211); the code uses CORR = 0.5.  The oracle follows the code,
    so its values are the asserted ones shifted by 0.5 - 0.355."""
    r = 0.1 * 0.1
    g0 = L.orc_filter_gauss(0.0, r)
    g1 = L.orc_filter_gauss(r, r)
    assert g0 - 0.5 == pytest.approx(1.2730000000000001 - 0.355, abs=1e-15)
    assert g1 - 0.5 == pytest.approx(0.6061526928041553 - 0.355, abs=1e-15)
    # the exact formula: ALPHA (1 - e_r / E_BETA) + CORR, e_r = 1 - exp(-BETA d / (2 r))
    e_r = 1.0 - math.exp(-1.953 * r / (r * 2.0))
    assert g1 == 0.918 * (1.0 - e_r / (1.0 - 0.14184788965323)) + 0.5


def test_filter_cone_cutoff():
    # beyond (1.1)^2 r the cone weight is zero (d2 > 1 branch)
    assert L.orc_filter_cone(1.2 * 1.2 * 0.01, 0.01) == 0.0


# ---------------------------------------------------------------------------
# algebra.rs:268, geometry.rs:217,227,233-242

def test_normalize_reference_value():
    out = np.zeros(3)
    assert L.orc_normalize(ptr(d3(1.0, -2.0, 3.0)), ptr(out)) == 1
    assert list(out) == [0.2672612419124244, -0.5345224838248488, 0.8017837257372732]
    assert L.orc_normalize(ptr(d3(0.0, 0.0, 0.0)), ptr(out)) == 0


def test_ray_new_and_target_reference_values():
    out = np.zeros(3)
    # Ray::new_from_elem(1,1,1,-1,-1,-1).target(2.0)
    assert L.orc_ray_target(ptr(d3(1, 1, 1)), ptr(d3(-1, -1, -1)), 2.0, ptr(out)) == 1
    assert list(out) == [-0.15470053837925168] * 3
    # Direction3::new_dir(1,1,1)
    assert L.orc_normalize(ptr(d3(1, 1, 1)), ptr(out)) == 1
    assert list(out) == [0.5773502691896258] * 3


def test_polygon_normal_reference_value():
    out = np.zeros(3)
    assert L.orc_polygon_normal(ptr(d3(0, 0, 0)), ptr(d3(2, 1, 0)), ptr(d3(0, 1, 2)), ptr(out)) == 1
    assert list(out) == [0.4082482904638631, -0.8164965809277261, 0.4082482904638631]


# ---------------------------------------------------------------------------
# physics.rs:372-377, 409-415

def test_color_normalize_and_wavelength_reference_values():
    out = np.zeros(3)
    L.orc_color_normalize(ptr(d3(0.4, 0.78, 1.0)), ptr(out))
    assert out[0] + out[1] + out[2] == 1.0
    assert list(out) == [0.1834862385321101, 0.35779816513761464, 0.4587155963302752]
    assert L.orc_decide_wavelength(ptr(out), 0.1) == 0
    assert L.orc_decide_wavelength(ptr(out), 0.3) == 1
    assert L.orc_decide_wavelength(ptr(out), 0.7) == 2
    L.orc_color_normalize(ptr(d3(0, 0, -1)), ptr(out))
    assert list(out) == [1 / 3, 1 / 3, 1 / 3]


def test_check_under_reference_values():
    ps = d3(0.1, 0.2, 0.3, 0.5, 0.8)
    for p, want in [(0.03, 0), (0.12, 1), (0.28, 2), (0.4, 3), (0.64, 4), (0.99, 5)]:
        assert L.orc_check_under(ptr(ps), 5, p) == want
    # inclusive: p == threshold stays in the bucket
    assert L.orc_check_under(ptr(ps), 5, 0.1) == 0


def test_relative_ior():
    # physics.rs:200-205 returns ior2 / ior1 (1.5); the reference test at
    # physics.rs:400 asserts 0.0, contradicting the code, and does not compile.
    assert L.orc_relative_ior(1.0, 1.5) == 1.5
    assert L.orc_relative_ior(0.0, 1.5) == 1.0


def test_schlick_endpoints():
    assert L.orc_schlick(1.0, 0.3) == 1.0
    assert L.orc_schlick(0.04, 1.0) == 0.04
    assert L.orc_schlick(0.04, 0.0) == 1.0


def test_specular_refraction_snell_and_tir():
    n = d3(0, 1, 0)
    for th in [0.1, 0.5, 1.0, 1.4]:
        v = d3(math.sin(th), -math.cos(th), 0.0)
        t = np.zeros(3)
        c2 = np.zeros(1)
        ok = L.orc_specular_refraction(ptr(n), ptr(v), 1.5, ptr(t), c2.ctypes.data_as(O.dp))
        assert ok == 1
        # Snell: sin(theta_t) = sin(theta_i) / eta; cos2 = (t . -n)
        assert abs(abs(t[0]) - math.sin(th) / 1.5) < 1e-12
        assert abs(c2[0] - (-t[1])) < 1e-12
    # total internal reflection from glass to air (eta = 1/1.5) beyond the critical angle
    v = d3(math.sin(1.0), -math.cos(1.0), 0.0)
    t = np.zeros(3)
    c2 = np.zeros(1)
    assert L.orc_specular_refraction(ptr(n), ptr(v), 1 / 1.5, ptr(t), c2.ctypes.data_as(O.dp)) == 0


# ---------------------------------------------------------------------------
# Philox4x32-10 known answers (Random123 kat_vectors)

def test_philox_known_answers():
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert O.philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert O.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_uniform_stream_range_and_moments():
    out = np.zeros(200000)
    L.orc_uniforms(1, 7, 3, 0x10000000, len(out), ptr(out))
    assert out.min() >= 0.0 and out.max() < 1.0
    assert np.all(out * 2 ** 24 == np.floor(out * 2 ** 24))  # 24-bit grid, exact in float
    assert abs(out.mean() - 0.5) < 5 * math.sqrt(1 / 12 / len(out))


def test_diffuse_reflection_is_uniform_hemisphere():
    """generate_random_dir_by_angle + flip (algebra.rs:225-235,
    surface.rs:408-417): uniform on the hemisphere, E[cos] = 1/2,
    E[cos^2] = 1/3."""
    n = d3(0.0, 0.0, 1.0) / 1.0
    k = 100000
    out = np.zeros(3 * k)
    L.orc_diffuse_samples(ptr(n), 5, k, ptr(out))
    d = out.reshape(k, 3)
    c = d @ n
    assert np.all(c > 0)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-12)
    assert abs(c.mean() - 0.5) < 5 * math.sqrt((1 / 3 - 0.25) / k)
    assert abs((c * c).mean() - 1 / 3) < 5 * math.sqrt((1 / 5 - 1 / 9) / k)


# ---------------------------------------------------------------------------
# radius schedule (util/iterator.rb:33-38)

@pytest.mark.parametrize("alpha", [0.5, 2.0 / 3.0])
def test_radius_schedule_closed_form(alpha):
    """r_i^2 / r_1^2 = prod_{k=1}^{i-1} (k+alpha)/(k+1)
    = Gamma(i+alpha) / (Gamma(1+alpha) Gamma(i+1))."""
    n = 512
    r2 = O.radius_schedule(0.005, alpha, 1, n)
    i = np.arange(1, n + 1)
    closed = np.exp(gammaln(i + alpha) - gammaln(1 + alpha) - gammaln(i + 1)) * 0.005 ** 2
    assert np.allclose(r2, closed, rtol=1e-12)
    # the Ruby driver's own recurrence on r (not r^2), iterator.rb:34-38
    r = 0.005
    for k in range(n):
        assert abs(r * r - r2[k]) <= 1e-15 * r2[k] * (k + 1)
        r = math.sqrt(((k + 1) + alpha) / ((k + 1) + 1.0)) * r
    # starting mid-schedule gives the same tail
    assert np.allclose(O.radius_schedule(0.005, alpha, 100, 10), r2[99:109], rtol=1e-14)
