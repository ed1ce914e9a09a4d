"""Scene and camera file loaders for the reference's file formats.

The reference's read_scene (scene.rs:20) and read_camera (camera.rs:108)
ignore their file argument and build fixed values; these loaders read the
files the project ships (example/*.scene, example/*.scr, doc/ebnf-camera.txt)
into the same types, so the GPU path runs the CLIs' actual inputs.

.scene: YAML with the sections
  light:    - type point (color, flux, position) | parallelogram (color,
              flux, position, dir1, dir2) | sun (..., ldir); 'power' is
              read as another spelling of 'flux', 'name' is ignored
  material: - type solid, name, emittance, reflectance, transmittance,
              specularrefl, ior, diffuseness, metalness, smoothness
  vertex:   - <alias>: [x, y, z]   (or a mapping <alias>: [x, y, z])
  object:   - type plain (normal, position) | sphere (center, radius) |
              polygon / parallelogram (pos1, pos2, pos3: vectors or vertex
              aliases), with name and material
.scr: "key : value" lines (# comments), keys with or without underscores.

Mapping onto the reference's types:
- a solid material is Material{emittance, transmittance, ior, Simple{
  reflectance, specularrefl, diffuseness, metalness, roughness}} with
  roughness = smoothness (the example files give 0.0 for mirror-like metals,
  so the field is the Simple surface's roughness parameter);
- a plain is Shape::Plain{nvec = normalize(normal), dist = -nvec.position}
  (scene.rs:361-381 writes the floor as nvec EY, dist 0 and the ceiling at
  y = 4 as nvec -EY, dist 4);
- a parallelogram or sun light's normal is normalize(dir1 x dir2)
  (scene.rs:23-30: dir1 = 1.33 EX, dir2 = 1.33 EZ, nvec = -EY); its colour is
  normalised (Color::normalize, physics.rs:61-71);
- polygons and parallelograms take dir1 = pos2 - pos1, dir2 = pos3 - pos1
  (geometry.rs:93-115).
"""
import math

import numpy as np
import yaml

from . import _abi as A
from . import scene as S


class SceneFileError(ValueError):
    pass


def _vec(v, where):
    if isinstance(v, (list, tuple)) and len(v) == 3:
        try:
            return tuple(float(x) for x in v)
        except (TypeError, ValueError):
            pass
    raise SceneFileError(f"{where}: expected [x, y, z], got {v!r}")


def _num(d, key, where, default=None):
    if key not in d:
        if default is None:
            raise SceneFileError(f"{where}: missing '{key}'")
        return default
    try:
        return float(d[key])
    except (TypeError, ValueError):
        raise SceneFileError(f"{where}: '{key}' is not a number: {d[key]!r}")


def _normalize(v):
    n = math.sqrt(sum(x * x for x in v))
    if n == 0.0:
        raise ValueError("zero-length vector")
    return tuple(x / n for x in v)


def read_scene(path, camera=None, walls_bounds_margin=0.01):
    """Parse a .scene file into a Scene (lights, materials, objects).  The
    camera is attached when given (read_camera's result)."""
    with open(path) as f:
        doc = yaml.safe_load(f)
    if not isinstance(doc, dict):
        raise SceneFileError(f"{path}: not a scene file")
    sc = S.Scene()
    # materials
    mats = {}
    for i, m in enumerate(doc.get("material") or []):
        where = f"{path}: material #{i + 1}"
        if not isinstance(m, dict) or m.get("type") != "solid":
            got = m.get("type") if isinstance(m, dict) else m
            raise SceneFileError(f"{where}: type must be 'solid', got {got!r}")
        name = m.get("name")
        if not name:
            raise SceneFileError(f"{where}: missing name")
        mat = S.simple(_vec(m.get("reflectance", [0, 0, 0]), where), _vec(m.get("specularrefl", [0, 0, 0]), where),
                       _num(m, "diffuseness", where, 0.0), _num(m, "metalness", where, 0.0),
                       _num(m, "smoothness", where, 0.0), emittance=_vec(m.get("emittance", [0, 0, 0]), where),
                       ior=_vec(m.get("ior", [0, 0, 0]), where),
                       transmittance=_vec(m.get("transmittance", [0, 0, 0]), where))
        mats[str(name)] = sc.add_material(mat)
    # vertex aliases
    verts = {}
    vsec = doc.get("vertex") or []
    if isinstance(vsec, dict):
        vsec = [{k: v} for k, v in vsec.items()]
    for i, v in enumerate(vsec):
        if not isinstance(v, dict) or len(v) != 1:
            raise SceneFileError(f"{path}: vertex #{i + 1}: expected '- name: [x, y, z]'")
        (k, val), = v.items()
        verts[str(k)] = _vec(val, f"{path}: vertex {k}")

    def pos(o, key, where):
        v = o.get(key)
        if isinstance(v, str):
            if v not in verts:
                raise SceneFileError(f"{where}: unknown vertex '{v}'")
            return verts[v]
        return _vec(v, f"{where}: {key}")

    # lights
    for i, l in enumerate(doc.get("light") or []):
        where = f"{path}: light #{i + 1}"
        if not isinstance(l, dict):
            raise SceneFileError(f"{where}: not a mapping")
        t = l.get("type")
        color = _vec(l.get("color", [1, 1, 1]), where)
        if "flux" not in l and "power" in l:
            l = dict(l, flux=l["power"])
        if t == "point":
            sc.lights.append(S.point_light(color, _num(l, "flux", where), _vec(l.get("position"), where)))
        elif t in ("parallelogram", "sun"):
            p0 = _vec(l.get("position"), where)
            d1, d2 = _vec(l.get("dir1"), where), _vec(l.get("dir2"), where)
            try:
                n = _normalize(np.cross(d1, d2))
            except ValueError:
                raise SceneFileError(f"{where}: dir1 and dir2 are parallel")
            if t == "parallelogram":
                sc.lights.append(S.parallelogram_light(color, _num(l, "flux", where), p0, n, d1, d2))
            else:
                try:
                    ldir = _normalize(_vec(l.get("ldir"), where))  # a Direction3 (light.rs:30-37)
                except ValueError:
                    raise SceneFileError(f"{where}: zero ldir")
                sc.lights.append(S.sun_light(color, _num(l, "flux", where), p0, n, d1, d2, ldir))
        else:
            raise SceneFileError(f"{where}: unknown light type {t!r}")
    # objects
    points = []
    for i, o in enumerate(doc.get("object") or []):
        where = f"{path}: object #{i + 1}"
        if not isinstance(o, dict):
            raise SceneFileError(f"{where}: not a mapping")
        t = o.get("type")
        mname = o.get("material")
        if mname not in mats:
            raise SceneFileError(f"{where}: unknown material {mname!r}")
        m = mats[mname]
        if t == "plain":
            try:
                n = _normalize(_vec(o.get("normal"), where))
            except ValueError:
                raise SceneFileError(f"{where}: zero normal")
            p = _vec(o.get("position"), where)
            sc.shapes.append(S.plain(n, -float(np.dot(n, p)), m))
            points.append(p)
        elif t == "sphere":
            c = _vec(o.get("center"), where)
            r = _num(o, "radius", where)
            sc.shapes.append(S.sphere(c, r, m))
            points += [tuple(np.add(c, r)), tuple(np.subtract(c, r))]
        elif t in ("polygon", "parallelogram"):
            p1, p2, p3 = pos(o, "pos1", where), pos(o, "pos2", where), pos(o, "pos3", where)
            try:
                sh = (S.polygon if t == "polygon" else S.parallelogram)(p1, p2, p3, m)
            except ValueError:
                raise SceneFileError(f"{where}: degenerate {t}")
            sc.shapes.append(sh)
            points += [p1, p2, p3]
            if t == "parallelogram":
                points.append(tuple(np.add(p2, np.subtract(p3, p1))))
        else:
            raise SceneFileError(f"{where}: unknown object type {t!r}")
    for l in sc.lights:
        points.append(l.pos)
    if camera is not None:
        sc.camera = camera
        points.append(camera.eye)
    if not points:
        raise SceneFileError(f"{path}: no objects")
    lo, hi = np.min(points, axis=0), np.max(points, axis=0)
    pad = walls_bounds_margin * max(1e-6, float(np.max(hi - lo)))
    sc.bounds_min, sc.bounds_max = tuple(lo - pad), tuple(hi + pad)
    return sc


# read_camera's defaults (camera.rs:109-128), in the file's key spelling
CAMERA_DEFAULTS = {
    "xresolution": 256, "yresolution": 256, "progressive": True, "antialias": True, "useclassic": True,
    "blur": True, "estimateradius": 0.2, "maxradiance": 0.01, "isosensitivity": 100.0, "shutterspeed": 0.004,
    "focallength": 50.0, "fnumber": 4.0, "focus": 7.0, "photonfilter": "none", "ambient": (0.0, 0.0, 0.0),
    "eyeposition": (1.0, 2.0, -4.5), "targetposition": (0.0, 1.0, 0.0), "upperdirection": (0.0, 1.0, 0.0),
    "nphoton": 100000, "samplephoton": 500,
}
_FILTERS = {"none": A.FILTER_NONE, "cone": A.FILTER_CONE, "gauss": A.FILTER_GAUSS}


def _parse_value(key, raw, where):
    raw = raw.strip()
    d = CAMERA_DEFAULTS.get(key)
    if isinstance(d, bool):
        if raw in ("yes", "true"):
            return True
        if raw in ("no", "false"):
            return False
        raise SceneFileError(f"{where}: '{key}' must be yes or no, got {raw!r}")
    if isinstance(d, int):
        try:
            return int(raw)
        except ValueError:
            raise SceneFileError(f"{where}: '{key}' must be an integer, got {raw!r}")
    if isinstance(d, float):
        try:
            return float(raw)
        except ValueError:
            raise SceneFileError(f"{where}: '{key}' must be a number, got {raw!r}")
    if isinstance(d, tuple):
        body = raw.strip()
        if not (body.startswith("[") and body.endswith("]")):
            raise SceneFileError(f"{where}: '{key}' must be [x, y, z], got {raw!r}")
        try:
            v = tuple(float(x) for x in body[1:-1].split(","))
        except ValueError:
            raise SceneFileError(f"{where}: '{key}' must be [x, y, z], got {raw!r}")
        if len(v) != 3:
            raise SceneFileError(f"{where}: '{key}' must have three components")
        return v
    if key == "photonfilter":
        if raw not in _FILTERS:
            raise SceneFileError(f"{where}: photon_filter must be none, cone or gauss, got {raw!r}")
        return raw
    return raw


def read_camera_file(path):
    """Parse a .scr file (doc/ebnf-camera.txt; keys with or without
    underscores) over read_camera's defaults.  Returns (Camera, settings)
    where settings holds every key, including nphoton, useclassic,
    estimateradius (not squared), maxradiance and photonfilter."""
    cfg = dict(CAMERA_DEFAULTS)
    with open(path) as f:
        for ln, line in enumerate(f, 1):
            text = line.split("#", 1)[0].strip()
            if not text:
                continue
            if ":" not in text:
                raise SceneFileError(f"{path}:{ln}: expected 'key : value'")
            k, v = text.split(":", 1)
            key = k.strip().replace("_", "")
            if key not in CAMERA_DEFAULTS:
                raise SceneFileError(f"{path}:{ln}: unknown camera parameter {k.strip()!r}")
            cfg[key] = _parse_value(key, v, f"{path}:{ln}")
    cam = S.Camera(int(cfg["xresolution"]), int(cfg["yresolution"]), cfg["eyeposition"], cfg["targetposition"],
                   cfg["upperdirection"], focus=float(cfg["focus"]), focal_len=float(cfg["focallength"]) / 1000.0,
                   f_number=float(cfg["fnumber"]), progressive=bool(cfg["progressive"]),
                   antialias=bool(cfg["antialias"]), blur=bool(cfg["blur"]), pfilter=_FILTERS[cfg["photonfilter"]])
    return cam, cfg
