"""CPU: image writers of rt / ppmpa (camera.rs:77-100, 212-214) and of
util/averager2.rb (PPM and OpenEXR of the mean over passes).

Pinned by closed forms: floor((r/max)^(1/2.2) * 255) at chosen points
(0, max/2 -> 186, max -> 255, above max -> 255), Ruby's Float#to_s and
Rust's Display on known values, the EXR byte layout (magic, version,
attribute list, offset table pointing at each line's y) and a round trip."""
import struct

import numpy as np
import pytest

from ppmpa_amd import image_io as I
from ppmpa_amd import photonmap_io as M


def test_radiance_to_rgb_points():
    m = 0.01
    img = np.array([0.0, m / 2, m, 3 * m, -1.0, m * 1e-9])
    # 0.5^(1/2.2) * 255 = 186.07...; (1e-9)^(1/2.2) * 255 = 0.020...; NaN -> 0 ('as i32')
    assert list(I.radiance_to_rgb(img, m)) == [0, 186, 255, 255, 0, 0]


def test_rgb_ppm_text():
    img = np.zeros((2, 3, 3))
    img[0, 1] = (0.02, 0.005, 0.0)
    txt = I.format_rgb_ppm(img, max_radiance=0.01)
    lines = txt.split("\n")
    assert lines[:5] == ["P3", "## max radiance = 0.01", "## image parameters = 1/250, F4, ISO100", "3 2", "255"]
    assert lines[5:11] == ["0 0 0", "255 186 0", "0 0 0", "0 0 0", "0 0 0", "0 0 0"]
    assert txt.endswith("\n") and len(lines) == 12


def test_pnm_header_slow_shutter():
    assert I.pnm_header(4, 4, max_radiance=2.5, shutter_speed=2.0, f_number=2.8, iso=400.0)[2] == \
        "## image parameters = 2, F2.8, ISO400"


@pytest.mark.parametrize("x,s", [(2.0, "2.0"), (0.01, "0.01"), (1e-4, "0.0001"), (1e-5, "1.0e-05"),
                                 (1.5e-7, "1.5e-07"), (1e15, "1000000000000000.0"), (1e16, "1.0e+16"),
                                 (0.1 + 0.2, "0.30000000000000004"), (0.0, "0.0"), (-3.0, "-3.0")])
def test_ruby_float(x, s):
    assert I.ruby_float(x) == s


def test_average_ppm():
    mean = np.zeros((1, 3, 3))
    mean[0, 0] = (4.0, 1.0, 0.0)     # one channel over max
    mean[0, 2] = (0.0, 0.0, 1e-12)   # lit in some pass: not the '0 0 0' of a never-lit pixel
    txt = I.format_average_ppm(mean, 2.0)
    lines = txt.strip().split("\n")
    assert lines[:4] == ["P3", "## max radiance = 2.0", "3 1", "255"]
    # (min(4/2,1))^(1/2.2)*255 = 255; (0.5)^(1/2.2)*255 = 186.07 -> 186 (to_i truncates)
    assert lines[4] == "255 186 0" and lines[5] == "0 0 0"
    v = int((5e-13) ** (1 / 2.2) * 255)
    assert lines[6] == f"0 0 {v}"


def test_exr_layout_and_round_trip():
    rng = np.random.default_rng(4)
    h, w = 5, 7
    mean = rng.uniform(0, 3, (h, w, 3))
    data = I.exr_bytes(mean, 2.0)
    assert struct.unpack_from("<ii", data, 0) == (20000630, 2)
    img, attrs = I.read_exr(data)
    assert attrs["compression"] == ("compression", b"\0")
    assert attrs["lineOrder"] == ("lineOrder", b"\0")
    assert struct.unpack("<iiii", attrs["dataWindow"][1]) == (0, 0, w - 1, h - 1)
    assert attrs["displayWindow"] == attrs["dataWindow"]
    assert struct.unpack("<f", attrs["pixelAspectRatio"][1]) == (1.0,)
    # channels sorted B, G, R as FLOAT, 16 bytes each, then the terminator
    ch = attrs["channels"][1]
    assert ch == b"".join(c + b"\0" + struct.pack("<iiii", 2, 0, 1, 1) for c in (b"B", b"G", b"R")) + b"\0"
    assert np.array_equal(img, (mean / 2.0).astype(np.float32))
    # the offset table points at each line's y and size
    hdr_end = data.index(b"screenWindowWidth") + len(b"screenWindowWidth\0float\0") + 4 + 4 + 1
    offs = struct.unpack_from(f"<{h}Q", data, hdr_end)
    for y, o in enumerate(offs):
        assert struct.unpack_from("<ii", data, o) == (y, 4 * w * 3)
    assert offs[-1] + 8 + 12 * w == len(data)


def test_write_exr_file(tmp_path):
    mean = np.ones((2, 2, 3))
    p = tmp_path / "x.exr"
    I.write_exr(str(p), mean, 4.0)
    img, _ = I.read_exr(p.read_bytes())
    assert np.all(img == 0.25)


def test_ppmf_round_trip():
    img = np.array([[[1.5e-3, 0.0, 2.0]], [[1e-300, 3.25, 7.0e10]]])
    txt = M.format_ppmf(img, max_radiance=0.05)
    back, mr = M.parse_ppmf(txt)
    assert mr == 0.05 and np.array_equal(back, img)
    assert txt.split("\n")[5] == "1.5e-3 0e0 2e0"
