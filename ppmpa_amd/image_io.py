"""Image outputs of the reference's pass and averaging tools.

- format_rgb_ppm: rt's non-progressive output, pnm_header then
  radiance_to_rgb per pixel (camera.rs:77-100, rt.rs:47-55);
- format_average_ppm / write_exr: util/averager2.rb's two outputs of the
  mean over passes (averager2.rb:84-110, 154-218): the PPM keeps 4 header
  lines and clips mean/max_radiance to [0, 1] before gamma 1/2.2 and x255
  (truncated); the EXR is single-part scanline OpenEXR, uncompressed, FLOAT
  channels B, G, R holding mean / max_radiance;
- read_exr: the inverse of write_exr (for tests and tools).
The per-pass .ppmf text (progressive mode) is photonmap_io.format_ppmf.
"""
import struct

import numpy as np

GAMMA = 1.0 / 2.2   # camera.rs:52, averager2.rb:11
RGBMAX = 255.0      # camera.rs:53


def radiance_to_rgb(img, max_radiance):
    """Camera::radiance_to_rgb (camera.rs:92-100): floor(min(r/max, 1)^gamma * 255);
    a negative component gives NaN there, which Rust's 'as i32' turns into 0."""
    d = np.minimum(np.asarray(img, np.float64) / max_radiance, 1.0)
    with np.errstate(invalid="ignore"):
        v = np.floor(np.power(d, GAMMA) * RGBMAX)
    return np.nan_to_num(v, nan=0.0).astype(np.int64)


def pnm_header(w, h, max_radiance=0.01, shutter_speed=0.004, f_number=4.0, iso=100.0):
    """Camera::pnm_header (camera.rs:77-90)."""
    from .photonmap_io import rust_f64
    ss = f"1/{rust_f64(1.0 / shutter_speed)}" if shutter_speed < 1.0 else rust_f64(shutter_speed)
    return ["P3", f"## max radiance = {rust_f64(max_radiance)}",
            f"## image parameters = {ss}, F{rust_f64(f_number)}, ISO{rust_f64(iso)}", f"{w} {h}", "255"]


def format_rgb_ppm(img, max_radiance=0.01, **kw):
    h, w, _ = img.shape
    rgb = radiance_to_rgb(img, max_radiance).reshape(-1, 3)
    return "\n".join(pnm_header(w, h, max_radiance, **kw) + [f"{a} {b} {c}" for a, b, c in rgb]) + "\n"


def ruby_float(x):
    """Ruby's Float#to_s, which averager2.rb uses for the max radiance line:
    shortest round-trip digits; fixed notation with at least one decimal
    for 1e-4 <= |x| < 1e16 (and 0.0), else d.ddde+XX / d.ddde-XX."""
    x = float(x)
    if np.isnan(x):
        return "NaN"
    if np.isinf(x):
        return "Infinity" if x > 0 else "-Infinity"
    if x == 0.0:
        return "-0.0" if np.signbit(x) else "0.0"
    mant, exp = np.format_float_scientific(x, unique=True, trim="-").split("e")
    e = int(exp)
    if -4 <= e < 16:
        s = np.format_float_positional(x, unique=True, trim="-")
        return s if "." in s else s + ".0"
    if "." not in mant:
        mant += ".0"
    return f"{mant}e{'+' if e >= 0 else '-'}{abs(e):02d}"


def format_average_ppm(mean, max_radiance):
    """averager2.rb's PPM of the mean over passes (clip: (min(c/n/max, 1)^gamma
    * 255).to_i, a pixel black in every pass stays '0 0 0')."""
    h, w, _ = mean.shape
    m = np.asarray(mean, np.float64).reshape(-1, 3)
    d = np.minimum(m / max_radiance, 1.0)
    v = np.trunc(np.power(np.maximum(d, 0.0), GAMMA) * RGBMAX).astype(np.int64)
    lines = ["P3", f"## max radiance = {ruby_float(max_radiance)}", f"{w} {h}", "255"]
    lines += ["0 0 0" if not np.any(px > 0) else f"{a} {b} {c}" for px, (a, b, c) in zip(m, v)]
    return "\n".join(lines) + "\n"


def _attr(name, typ, payload):
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(payload)) + payload


def exr_bytes(mean, max_radiance):
    """averager2.rb's output_exr: header (channels B, G, R as FLOAT;
    compression NONE; data and display window (0, 0, w-1, h-1); lineOrder
    INCREASING_Y; pixelAspectRatio 1; screenWindowCenter (0, 0);
    screenWindowWidth 1), the line offset table, then per line
    y, data size, B values, G values, R values (little-endian float32)."""
    h, w, _ = mean.shape
    ch = b"".join(c.encode() + b"\0" + struct.pack("<iiii", 2, 0, 1, 1) for c in "BGR") + b"\0"
    box = struct.pack("<iiii", 0, 0, w - 1, h - 1)
    header = struct.pack("<ii", 20000630, 2)
    header += _attr("channels", "chlist", ch)
    header += _attr("compression", "compression", b"\0")
    header += _attr("dataWindow", "box2i", box)
    header += _attr("displayWindow", "box2i", box)
    header += _attr("lineOrder", "lineOrder", b"\0")
    header += _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    header += _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0))
    header += _attr("screenWindowWidth", "float", struct.pack("<f", 1.0))
    header += b"\0"
    datalen = 4 * w * 3
    linelen = 8 + datalen
    offset = len(header) + 8 * h
    table = b"".join(struct.pack("<Q", offset + linelen * y) for y in range(h))
    v = (np.asarray(mean, np.float64) / max_radiance).astype("<f4")
    lines = b"".join(struct.pack("<ii", y, datalen) + v[y, :, 2].tobytes() + v[y, :, 1].tobytes() +
                     v[y, :, 0].tobytes() for y in range(h))
    return header + table + lines


def write_exr(path, mean, max_radiance):
    with open(path, "wb") as f:
        f.write(exr_bytes(mean, max_radiance))


def read_exr(data):
    """(image[h, w, 3] of float32 R, G, B; header attributes) of an
    uncompressed single-part scanline EXR with FLOAT channels."""
    if struct.unpack_from("<i", data, 0)[0] != 20000630:
        raise ValueError("not an OpenEXR file")
    pos = 8
    attrs = {}
    while data[pos] != 0:
        e = data.index(b"\0", pos)
        name = data[pos:e].decode()
        e2 = data.index(b"\0", e + 1)
        typ = data[e + 1:e2].decode()
        size = struct.unpack_from("<i", data, e2 + 1)[0]
        attrs[name] = (typ, data[e2 + 5:e2 + 5 + size])
        pos = e2 + 5 + size
    pos += 1
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    chl = attrs["channels"][1]
    names, p = [], 0
    while chl[p] != 0:
        e = chl.index(b"\0", p)
        names.append(chl[p:e].decode())
        if struct.unpack_from("<i", chl, e + 1)[0] != 2:
            raise ValueError("only FLOAT channels are read")
        p = e + 17
    if attrs["compression"][1] != b"\0":
        raise ValueError("only uncompressed files are read")
    offsets = struct.unpack_from(f"<{h}Q", data, pos)
    img = np.zeros((h, w, 3), np.float32)
    for y in range(h):
        o = offsets[y]
        yy, n = struct.unpack_from("<ii", data, o)
        vals = np.frombuffer(data, "<f4", count=w * len(names), offset=o + 8).reshape(len(names), w)
        for k, c in enumerate(names):
            img[yy - y0, :, "RGB".index(c)] = vals[k]
    return img, attrs
