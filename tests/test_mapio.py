"""The pm / rt photon-map text format and the per-pass image text
(pm.rs:44-45,62-75, photonmap.rs:31-74, camera.rs:77-90,212-214).  CPU only."""
import numpy as np
import pytest

from ppmpa_amd import photonmap_io as M


def test_rust_display_formatting():
    # Rust `{}` for f64: shortest round trip, no exponent, no trailing ".0"
    assert M.rust_f64(1.0) == "1"
    assert M.rust_f64(-2.0) == "-2"
    assert M.rust_f64(0.1) == "0.1"
    assert M.rust_f64(1.1 * 3.0) == "3.3000000000000003"  # algebra.rs:275
    assert M.rust_f64(1e-7) == "0.0000001"
    assert M.rust_f64(5.0 / 200000) == "0.000025"
    assert M.rust_f64(-0.0) == "-0"
    # Rust `{:e}`
    assert M.rust_f64_exp(1.0) == "1e0"
    assert M.rust_f64_exp(0.15) == "1.5e-1"
    assert M.rust_f64_exp(123.456) == "1.23456e2"
    assert M.rust_f64_exp(0.0) == "0e0"


def test_format_parse_round_trip():
    rng = np.random.default_rng(1)
    n = 500
    pos = rng.uniform(-2, 2, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    wl = rng.integers(0, 3, n).astype(np.uint8)
    text = M.format_map(100000, 5.0 / 100000, pos, d, wl)
    lines = text.split("\n")
    assert lines[0] == "100000" and lines[1] == "0.00005"
    assert lines[2].split(" ")[0] in ("Red", "Green", "Blue") and len(lines[2].split(" ")) == 7
    m = M.parse_map(text)
    assert m.n_emitted == 100000 and m.power == 5e-5
    assert np.array_equal(m.pos, pos) and np.array_equal(m.dir, d) and np.array_equal(m.wl, wl)


def test_reader_quirks():
    # power line that does not parse -> 1.0 (photonmap.rs:37-45)
    m = M.parse_map("12\nnot-a-number\nGreen 1 2 3 0 1 0\nPurple 0 0 0 1 0 0\n")
    assert m.power == 1.0
    assert list(m.wl) == [1, 0]  # an unknown channel name reads as Red
    # extra fields are ignored
    m = M.parse_map("1\n2\nBlue 1 2 3 0 0 1 extra\n")
    assert list(m.wl) == [2] and m.power == 2.0
    # empty body
    assert len(M.parse_map("0\n1\n").wl) == 0


@pytest.mark.parametrize("bad", ["1\n1\nRed 1 2 3 0 1\n", "1\n1\nRed 1 2 x 0 1 0\n", "1\n1\n\nRed 1 2 3 0 1 0\n",
                                 "1\n1\nRed 1 2 3 0 0 0\n"])
def test_reader_rejects_malformed_lines(bad):
    # the reference panics (unwrap on parse / index / zero direction)
    with pytest.raises(M.PhotonMapParseError):
        M.parse_map(bad)


def test_ppmf_round_trip():
    img = np.random.default_rng(2).uniform(0, 3, (4, 5, 3))
    text = M.format_ppmf(img)
    lines = text.split("\n")
    assert lines[0] == "P3" and lines[1] == "## max radiance = 0.01"
    assert lines[2] == "## image parameters = 1/250, F4, ISO100"
    assert lines[3] == "5 4" and lines[4] == "255"
    back, maxrad = M.parse_ppmf(text)
    assert maxrad == 0.01 and np.array_equal(back, img)
