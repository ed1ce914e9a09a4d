"""Photon-map text format of the reference's `pm | rt` pipeline.

Writer: pm.rs:44-45 (line 1 = number of photons emitted, line 2 = photon
power) and pm.rs:62-75 (one "<Red|Green|Blue> px py pz dx dy dz" line per
stored photon, Rust `{}` f64 formatting).  Reader: read_map
(photonmap.rs:31-74): line 1 ignored, line 2 parsed as the power (1.0 when it
does not parse), then one photon per line; an unknown channel name reads as
Red, a malformed number is an error (the reference panics).

Also the per-pass radiance image (.ppmf): pnm_header (camera.rs:77-90)
plus one `{:e} {:e} {:e}` line per pixel (camera.rs:212-214).
"""
from dataclasses import dataclass

import numpy as np

WL_NAMES = ("Red", "Green", "Blue")


class PhotonMapParseError(ValueError):
    pass


def rust_f64(x: float) -> str:
    """Rust's Display for f64: shortest round-trip digits, no exponent."""
    x = float(x)
    if np.isnan(x):
        return "NaN"
    if np.isinf(x):
        return "inf" if x > 0 else "-inf"
    return np.format_float_positional(x, unique=True, trim="-")


def rust_f64_exp(x: float) -> str:
    """Rust's LowerExp ({:e}) for f64: shortest digits, e.g. 1.5e-1, 1e0."""
    x = float(x)
    if np.isnan(x):
        return "NaN"
    if np.isinf(x):
        return "inf" if x > 0 else "-inf"
    s = np.format_float_scientific(x, unique=True, trim="-")
    mant, exp = s.split("e")
    return f"{mant}e{int(exp)}"


@dataclass
class MapText:
    n_emitted: int
    power: float
    pos: np.ndarray
    dir: np.ndarray
    wl: np.ndarray


def parse_map(path_or_text) -> MapText:
    """read_map (photonmap.rs:31-74).  Directions are returned as written;
    the library re-normalises them on upload (Ray::new_from_elem)."""
    if "\n" in path_or_text or not _is_file(path_or_text):
        text = path_or_text
    else:
        with open(path_or_text) as f:
            text = f.read()
    lines = text.split("\n")
    line1 = lines[0] if lines else ""
    try:
        n_emitted = int(line1.strip())
    except ValueError:
        n_emitted = 0
    line2 = lines[1] if len(lines) > 1 else ""
    try:
        power = float(line2.strip())
    except ValueError:
        power = 1.0
    body = lines[2:]
    if body and body[-1] == "":
        body = body[:-1]  # str::lines drops the final terminator
    n = len(body)
    pos = np.zeros((n, 3))
    dir = np.zeros((n, 3))
    wl = np.zeros(n, np.uint8)
    for i, line in enumerate(body):
        if line.endswith("\r"):
            line = line[:-1]
        el = line.split(" ")
        if len(el) < 7:
            raise PhotonMapParseError(f"photon line {i + 3}: expected 7 fields, got {len(el)}")
        wl[i] = WL_NAMES.index(el[0]) if el[0] in WL_NAMES else 0
        try:
            vals = [float(e) for e in el[1:7]]
        except ValueError as e:
            raise PhotonMapParseError(f"photon line {i + 3}: {e}") from None
        pos[i] = vals[:3]
        dir[i] = vals[3:6]
        if np.all(dir[i] == 0.0):
            raise PhotonMapParseError(f"photon line {i + 3}: zero direction")
    return MapText(n_emitted, power, pos, dir, wl)


def _is_file(s):
    import os
    return len(s) < 4096 and os.path.isfile(s)


def format_map(n_emitted, power, pos, dir, wl) -> str:
    """pm's stdout (pm.rs:44-45,62-75)."""
    out = [str(int(n_emitted)), rust_f64(power)]
    for p, d, w in zip(np.asarray(pos), np.asarray(dir), np.asarray(wl)):
        out.append(" ".join([WL_NAMES[int(w)]] + [rust_f64(v) for v in p] + [rust_f64(v) for v in d]))
    return "\n".join(out) + "\n"


def write_map(path, n_emitted, power, pos, dir, wl):
    with open(path, "w") as f:
        f.write(format_map(n_emitted, power, pos, dir, wl))


def format_ppmf(img, max_radiance=0.01, shutter_speed=0.004, f_number=4.0, iso=100.0) -> str:
    """A per-pass radiance image as ppmpa/rt print it in progressive mode:
    pnm_header (camera.rs:77-90) then one radiance_to_string line per pixel
    (camera.rs:212-214, '{:e} {:e} {:e}'; rt.rs:47-55)."""
    h, w, _ = img.shape
    ss = f"1/{rust_f64(1.0 / shutter_speed)}" if shutter_speed < 1.0 else rust_f64(shutter_speed)
    head = ["P3", f"## max radiance = {rust_f64(max_radiance)}",
            f"## image parameters = {ss}, F{rust_f64(f_number)}, ISO{rust_f64(iso)}", f"{w} {h}", "255"]
    body = [" ".join(rust_f64_exp(v) for v in px) for px in img.reshape(-1, 3)]
    return "\n".join(head + body) + "\n"


def parse_ppmf(text):
    """Inverse of format_ppmf: (image[h, w, 3], max_radiance)."""
    lines = text.split("\n")
    maxrad = float(lines[1].split("=")[1])
    w, h = (int(x) for x in lines[3].split())
    vals = np.array([[float(v) for v in lines[5 + i].split()] for i in range(w * h)])
    return vals.reshape(h, w, 3), maxrad
