"""Command lines of the reference's tools, on the GPU path.

  python -m ppmpa_amd.pm [-h] <scene file> [<#photon>]          pm.rs:21-50
      photon pass: the photon map as text on stdout (pm.rs:44-45,62-75)
  python -m ppmpa_amd.rt [-h] <scene file> <camera file> [<radius>]   rt.rs:20-59
      eye pass over the photon map read from stdin (read_map,
      photonmap.rs:31-74): pnm header then radiance (progressive camera) or
      RGB lines (rt.rs:47-55)
  python -m ppmpa_amd.rtc [-h] <screen file> <scene file>             rtc.rs:17-47
      the classic ray tracer (trace_ray_classic, tracer.rs:221-259), no
      photons: pnm header then radiance (progressive camera) or RGB lines
  python -m ppmpa_amd.ppmpa [-nc|-h] <#photon> <radius> <camera file> <scene file>
      one in-process pass (ppmpa.rs:16-41): header + radiance lines
  python -m ppmpa_amd.iterate [-ppm|-exr <file>] <#photon> <#iteration> <radius> <screen> <scene>
      a PPM-PA run (util/iterator.rb, alpha 0.5) averaged as util/averager2.rb
      does, written as OpenEXR (default) or PPM

Differences from the reference, all deliberate: the scene and camera files
are read (the reference ignores them and uses its built-in scene,
scene.rs:20, camera.rs:108); the random numbers come from the counter-based
Philox streams keyed by --seed and --pass (the reference's thread_rng is
unseeded); iterate runs its passes in one device session instead of
spawning pm | rt processes.
"""
import argparse
import sys

import numpy as np

from . import image_io as I
from . import loaders as L
from . import photonmap_io as M
from . import tracer as T

MAX_TRACE = 10          # tracer.rs:27
DEF_NPHOTON = 100_000   # pm.rs:17, ppmpa.rs:17
DEF_RADIUS = 0.1        # rt.rs:18, ppmpa.rs:18
ITER_ALPHA = 0.5        # util/iterator.rb:17


def _ctx():
    return T.Context(0)


def pm_main(argv=None):
    ap = argparse.ArgumentParser(prog="pm", description="photon pass: output the photon map to stdout")
    ap.add_argument("scene")
    ap.add_argument("nphoton", nargs="?", default=str(DEF_NPHOTON))
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--pass", dest="pass_", type=int, default=1)
    a = ap.parse_args(argv)
    try:
        n = int(a.nphoton)
    except ValueError:
        n = DEF_NPHOTON  # pm.rs:30-35
    sc = L.read_scene(a.scene)
    from . import scene as S
    sc.camera = S.Camera(8, 8, (0.0, 0.0, -1.0), (0.0, 0.0, 0.0))  # the photon pass has no camera
    ctx = _ctx()
    ds = T.DeviceScene(ctx, sc)
    ph = T.trace_photons(ctx, ds, a.seed, a.pass_, n, use_classic=True, max_depth=MAX_TRACE)
    pos, dir, wl, _ = ph.download()
    sys.stdout.write(M.format_map(n, T.photon_power(sc, n), pos, dir, wl))
    return 0


def _camera_scene(scene_file, camera_file):
    cam, cfg = L.read_camera_file(camera_file)
    sc = L.read_scene(scene_file, camera=cam)
    return sc, cam, cfg


def _header_and_body(img, cam, cfg):
    w = img.shape[1]
    h = img.shape[0]
    kw = dict(max_radiance=cfg["maxradiance"], shutter_speed=cfg["shutterspeed"], f_number=cfg["fnumber"],
              iso=cfg["isosensitivity"])
    if cam.progressive:
        return M.format_ppmf(img, **kw)
    return I.format_rgb_ppm(img, **kw)


def rt_main(argv=None):
    ap = argparse.ArgumentParser(prog="rt", description="eye pass over the photon map read from stdin")
    ap.add_argument("scene")
    ap.add_argument("camera")
    ap.add_argument("radius", nargs="?", default=str(DEF_RADIUS))
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--pass", dest="pass_", type=int, default=1)
    a = ap.parse_args(argv)
    try:
        r = float(a.radius)
    except ValueError:
        r = DEF_RADIUS  # rt.rs:28-36
    sc, cam, cfg = _camera_scene(a.scene, a.camera)
    ctx = _ctx()
    ds = T.DeviceScene(ctx, sc)
    pmap, _ = T.read_map(ctx, sys.stdin.read())
    img = T.trace_eye(ctx, ds, pmap, a.seed, a.pass_, r * r, use_classic=True, max_depth=MAX_TRACE,
                      pfilter=cam.pfilter)
    sys.stdout.write(_header_and_body(img, cam, cfg))
    return 0


def rtc_main(argv=None):
    ap = argparse.ArgumentParser(prog="rtc", description="classic ray tracer: direct light + ambient, no photon map")
    ap.add_argument("screen")
    ap.add_argument("scene")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--pass", dest="pass_", type=int, default=1)
    a = ap.parse_args(argv)
    sc, cam, cfg = _camera_scene(a.scene, a.screen)
    ctx = _ctx()
    ds = T.DeviceScene(ctx, sc)
    img = T.trace_classic(ctx, ds, a.seed, a.pass_, cfg["ambient"])
    sys.stdout.write(_header_and_body(img, cam, cfg))  # rtc.rs:25-27,32-39
    return 0


def ppmpa_main(argv=None):
    ap = argparse.ArgumentParser(prog="ppmpa", description="one in-process PPM pass")
    ap.add_argument("-nc", action="store_true", help="no classic direct light (ppmpa.rs:48-52)")
    ap.add_argument("nphoton")
    ap.add_argument("radius")
    ap.add_argument("camera")
    ap.add_argument("scene")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--pass", dest="pass_", type=int, default=1)
    a = ap.parse_args(argv)
    try:
        n = int(a.nphoton)
    except ValueError:
        n = DEF_NPHOTON
    try:
        r = float(a.radius)
    except ValueError:
        r = DEF_RADIUS
    sc, cam, cfg = _camera_scene(a.scene, a.camera)
    ctx = _ctx()
    ds = T.DeviceScene(ctx, sc)
    uc = not a.nc
    ph = T.trace_photons(ctx, ds, a.seed, a.pass_, n, use_classic=uc, max_depth=MAX_TRACE)
    pmap = T.build_photonmap(ctx, ph, T.photon_power(sc, n), ds)
    img = T.trace_eye(ctx, ds, pmap, a.seed, a.pass_, r * r, use_classic=uc, max_depth=MAX_TRACE,
                      pfilter=cam.pfilter)
    kw = dict(max_radiance=cfg["maxradiance"], shutter_speed=cfg["shutterspeed"], f_number=cfg["fnumber"],
              iso=cfg["isosensitivity"])
    sys.stdout.write(M.format_ppmf(img, **kw))  # ppmpa.rs:35-40 always prints radiance
    return 0


def iterate_main(argv=None):
    ap = argparse.ArgumentParser(prog="iterate", description="PPM-PA passes averaged into one image")
    fmt = ap.add_mutually_exclusive_group()
    fmt.add_argument("-ppm", metavar="FILE", help="write the averaged image as PPM")
    fmt.add_argument("-exr", metavar="FILE", help="write the averaged image as OpenEXR (the default format)")
    ap.add_argument("nphoton", type=int)
    ap.add_argument("niterate", type=int)
    ap.add_argument("radius", type=float)
    ap.add_argument("screen")
    ap.add_argument("scene")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--alpha", type=float, default=ITER_ALPHA)
    a = ap.parse_args(argv)
    sc, cam, cfg = _camera_scene(a.scene, a.screen)
    ctx = _ctx()
    ds = T.DeviceScene(ctx, sc)
    c = T.pass_cfg(seed=a.seed, n_photon=a.nphoton, radius0=a.radius, alpha=a.alpha, max_depth=MAX_TRACE,
                   use_classic=bool(cfg["useclassic"]), pfilter=cam.pfilter)
    s = T.Session(ctx, ds, c)
    s.run(a.niterate)
    mean, _, n = s.image(with_var=False)
    s.close()
    maxrad = cfg["maxradiance"]
    if a.ppm:
        with open(a.ppm, "w") as f:
            f.write(I.format_average_ppm(mean, maxrad))
    else:
        data = I.exr_bytes(mean, maxrad)
        if a.exr:
            with open(a.exr, "wb") as f:
                f.write(data)
        else:
            sys.stdout.buffer.write(data)
    print(f"{n} passes, {a.nphoton} photons each", file=sys.stderr)
    return 0
