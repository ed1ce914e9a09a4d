"""Python mirror of the reference's tracer / photon-map call surface over the
gfx950 C ABI (libppm.so).

  reference                                     here
  ---------------------------------------------  ------------------------------------
  Light::generate_photon + trace_photon           trace_photons()          (tracer.rs:31, light.rs:67)
    (ppmpa.rs:86-92, pm.rs:56-75)
  build_photonmap (photonmap.rs:23-29)           build_photonmap()
  read_map (photonmap.rs:31-74)                  read_map()  (text parsing in photonmap_io)
  KdTree::within (tracer.rs:180)                 PhotonMap.within()
  estimate_radiance (tracer.rs:179-195)          PhotonMap.estimate_radiance()
  trace_ray per pixel (ppmpa.rs:81-82)         trace_eye()
  util/iterator.rb + util/averager2.rb           Session / run_passes()

Every call goes to the HIP library; there is no CPU fallback.  Errors raise
PpmError carrying the library's status and message.
"""
import atexit
import ctypes as C
import weakref

import numpy as np

from . import _abi as A

# Contexts alive in this process.  At exit they are closed while the
# interpreter is still whole, each after every object made on it (the
# objects' destroy calls wait on the context's stream, which ppm_destroy
# frees); left to the interpreter's shutdown the order would be arbitrary.
_CONTEXTS = weakref.WeakSet()


@atexit.register
def _close_all():
    for c in list(_CONTEXTS):
        try:
            c.close()
        except Exception:
            pass


class PpmError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"ppm status {status}: {msg}")
        self.status = status


def _dp(a):
    return a.ctypes.data_as(A.dp)


class Context:
    """One HIP device + stream (ppm_create)."""

    def __init__(self, device=0):
        self.lib = A.load()
        self.h = None
        self._objs = weakref.WeakSet()  # everything made on this context
        h = C.c_void_p()
        s = self.lib.ppm_create(C.byref(h), int(device), None)
        if s != 0:
            raise PpmError(s, self.lib.ppm_status_string(s).decode())
        self.h = h
        _CONTEXTS.add(self)

    def _own(self, obj):
        """Record an object made on this context (closed before it)."""
        self._objs.add(obj)

    def check(self, s):
        if s != 0:
            msg = self.lib.ppm_last_error(self.h).decode() or self.lib.ppm_status_string(s).decode()
            raise PpmError(s, msg)

    def sync(self):
        self.check(self.lib.ppm_sync(self.h))

    def close(self):
        """ppm_destroy, after closing every object still open on the context
        (batches and sessions first, scenes last)."""
        if self.h:
            rank = {Batch: 0, Session: 1, PhotonMap: 2, PhotonSet: 3, DeviceScene: 4}
            for o in sorted(list(self._objs), key=lambda o: rank.get(type(o), 0)):
                o.close()
            self.lib.ppm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceScene:
    """read_scene + read_camera uploaded (ppm_scene_create)."""

    def __init__(self, ctx, scene):
        self.ctx, self.scene = ctx, scene
        desc, keep = scene.desc()
        h = C.c_void_p()
        ctx.check(ctx.lib.ppm_scene_create(ctx.h, C.byref(desc), C.byref(h)))
        self.h = h
        self._keep = keep
        ctx._own(self)

    @property
    def pixels(self):
        return self.scene.camera.xreso * self.scene.camera.yreso

    def set_params(self, albedo=None, survival=None, light_scale=1.0):
        """Reflectance c1 of materials 0..n-1 (n x 3), the survival
        probabilities q of their diffuse bounce (None: q = c1) and a factor on
        every light flux and emittance, relative to the scene as created
        (ppm_scene_set_params).  Sessions created afterwards see them."""
        a = None if albedo is None else np.ascontiguousarray(albedo, np.float64).reshape(-1, 3)
        q = None if survival is None else np.ascontiguousarray(survival, np.float64).reshape(-1, 3)
        n = len(a) if a is not None else (len(q) if q is not None else 0)
        if a is not None and q is not None and len(q) != len(a):
            raise ValueError("albedo and survival need the same number of materials")
        self.ctx.check(self.ctx.lib.ppm_scene_set_params(self.ctx.h, self.h, n, _dp(a) if a is not None else None,
                                                         _dp(q) if q is not None else None, float(light_scale)))

    def stats(self):
        """(shapes tested one by one, shapes in the BVH, BVH leaf level)."""
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        self.ctx.check(self.ctx.lib.ppm_scene_stats(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def intersect(self, origins, dirs):
        """calc_intersection (tracer.rs:306-335) for n rays: (t, shape index),
        (inf, -1) where nothing is hit."""
        o = np.ascontiguousarray(origins, np.float64).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
        n = len(o)
        t = np.empty(n, np.float64)
        obj = np.empty(n, np.int32)
        self.ctx.check(self.ctx.lib.ppm_intersect(self.ctx.h, self.h, _dp(o), _dp(d), n, _dp(t),
                                                  obj.ctypes.data_as(C.POINTER(C.c_int32))))
        return t, obj

    def close(self):
        if self.h:
            self.ctx.lib.ppm_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PhotonSet:
    """A device-resident photon set (ppm_photons_*)."""

    def __init__(self, ctx, capacity):
        self.ctx = ctx
        self.capacity = int(capacity)
        h = C.c_void_p()
        ctx.check(ctx.lib.ppm_photons_create(ctx.h, self.capacity, C.byref(h)))
        self.h = h
        self.power = 1.0
        ctx._own(self)

    def count(self):
        n = C.c_uint64()
        self.ctx.check(self.ctx.lib.ppm_photons_count(self.ctx.h, self.h, C.byref(n)))
        return n.value

    def upload(self, pos, dir, wl):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        dir = np.ascontiguousarray(dir, dtype=np.float64).reshape(-1, 3)
        wl = np.ascontiguousarray(wl, dtype=np.uint8).reshape(-1)
        n = pos.shape[0]
        if dir.shape[0] != n or wl.shape[0] != n:
            raise ValueError("pos, dir and wl lengths differ")
        self.ctx.check(self.ctx.lib.ppm_photons_upload(self.ctx.h, self.h, n, _dp(pos), _dp(dir),
                                                       wl.ctypes.data_as(A.u8p)))

    def download(self):
        """(pos[n,3], dir[n,3], wl[n], meta[n]) widened to double."""
        n = self.count()
        pos = np.zeros((max(n, 1), 3))
        dir = np.zeros((max(n, 1), 3))
        wl = np.zeros(max(n, 1), np.uint8)
        meta = np.zeros(max(n, 1), np.uint32)
        got = C.c_uint64()
        self.ctx.check(self.ctx.lib.ppm_photons_download(self.ctx.h, self.h, n, _dp(pos), _dp(dir),
                                                         wl.ctypes.data_as(A.u8p), meta.ctypes.data_as(A.u32p),
                                                         C.byref(got)))
        n = got.value
        return pos[:n], dir[:n], wl[:n], meta[:n]

    def close(self):
        if self.h:
            self.ctx.lib.ppm_photons_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def photon_power(scene, n_photon):
    """ppmpa.rs:30 / pm.rs:40-41: Σ flux / nphoton."""
    return sum(l.flux for l in scene.lights) / float(n_photon)


def trace_photons(ctx, dscene, seed, pass_, n_photon, use_classic=True, max_depth=10, out=None):
    """Σ_lights Σ_i trace_photon(uc, M_AIR, objs, 0, light.generate_photon())
    for one pass; returns the PhotonSet (ppm_trace_photons)."""
    if out is None:
        # one region per tracing wave, each sized for its photons' deepest walks
        out = PhotonSet(ctx, (int(n_photon) + 8192) * int(max_depth))
    ctx.check(ctx.lib.ppm_trace_photons(ctx.h, dscene.h, int(seed), int(pass_), int(n_photon), int(use_classic),
                                        int(max_depth), out.h))
    out.power = photon_power(dscene.scene, n_photon)
    return out


class PhotonMap:
    """build_photonmap's PhotonMap (photonmap.rs:16-29) on the device."""

    def __init__(self, ctx, capacity):
        self.ctx = ctx
        self.capacity = int(capacity)
        h = C.c_void_p()
        ctx.check(ctx.lib.ppm_map_create(ctx.h, self.capacity, C.byref(h)))
        self.h = h
        self.power = 1.0
        ctx._own(self)

    def build(self, photons, power, dscene=None):
        self.power = float(power)
        self.ctx.check(self.ctx.lib.ppm_map_build(self.ctx.h, dscene.h if dscene is not None else None, photons.h,
                                                  self.power, self.h))
        return self

    def build_grid(self, photons, power, r2, dscene=None):
        """The map as the PPM-PA passes build it for one radius (hashed cells
        of side >= 2.1 sqrt(r2)): within() and estimate_radiance() then run
        the passes' own cell walk and shade kernel, for any r2 up to this
        one (ppm_map_build_grid)."""
        self.power = float(power)
        self.ctx.check(self.ctx.lib.ppm_map_build_grid(self.ctx.h, dscene.h if dscene is not None else None,
                                                       photons.h, self.power, float(r2), self.h))
        return self

    def within(self, q, r2, refine64=False):
        """KdTree::within(q, r2, squared_euclidean) for each query point:
        a list of ascending index arrays."""
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
        nq = q.shape[0]
        off = np.zeros(nq + 1, np.uint64)
        cap = 0
        idx = np.zeros(1, np.uint32)
        s = self.ctx.lib.ppm_radius_query(self.ctx.h, self.h, _dp(q), nq, float(r2), int(refine64), cap,
                                          idx.ctypes.data_as(A.u32p), off.ctypes.data_as(A.u64p))
        if s == A.PPM_E_RANGE:
            cap = int(off[nq])
            idx = np.zeros(max(cap, 1), np.uint32)
            s = self.ctx.lib.ppm_radius_query(self.ctx.h, self.h, _dp(q), nq, float(r2), int(refine64), cap,
                                              idx.ctypes.data_as(A.u32p), off.ctypes.data_as(A.u64p))
        self.ctx.check(s)
        return [idx[int(off[i]):int(off[i + 1])] for i in range(nq)]

    def estimate_radiance(self, q, nrm, r2, pfilter=A.FILTER_NONE, refine64=False):
        """estimate_radiance (tracer.rs:179-195) at (position, normal) pairs."""
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
        nrm = np.ascontiguousarray(nrm, dtype=np.float64).reshape(-1, 3)
        out = np.zeros_like(q)
        self.ctx.check(self.ctx.lib.ppm_gather(self.ctx.h, self.h, _dp(q), _dp(nrm), q.shape[0], float(r2),
                                               int(pfilter), int(refine64), _dp(out)))
        return out

    def close(self):
        if self.h:
            self.ctx.lib.ppm_map_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_photonmap(ctx, photons, power, dscene=None, capacity=None):
    """build_photonmap(pw, radius, phs, nsample) (photonmap.rs:23-29)."""
    m = PhotonMap(ctx, capacity or photons.capacity)
    return m.build(photons, power, dscene)


def read_map(ctx, path_or_text):
    """read_map (photonmap.rs:31-74): parse the pm text format and load the
    photons; returns (PhotonMap, PhotonSet)."""
    from .photonmap_io import parse_map
    pm = parse_map(path_or_text)
    ph = PhotonSet(ctx, max(1, len(pm.wl)))
    ph.upload(pm.pos, pm.dir, pm.wl)
    ph.power = pm.power
    m = PhotonMap(ctx, ph.capacity).build(ph, pm.power)
    return m, ph


def trace_eye(ctx, dscene, pmap, seed, pass_, r2, use_classic=True, max_depth=10, pfilter=A.FILTER_NONE):
    """One eye pass: trace_ray for every pixel (ppmpa.rs:81-82); (H, W, 3)."""
    cam = dscene.scene.camera
    out = np.zeros((cam.yreso, cam.xreso, 3))
    ctx.check(ctx.lib.ppm_trace_eye(ctx.h, dscene.h, pmap.h, int(seed), int(pass_), float(r2), int(use_classic),
                                    int(max_depth), int(pfilter), _dp(out)))
    return out


def trace_classic(ctx, dscene, seed, pass_, ambient=(0.0, 0.0, 0.0)):
    """rtc: trace_ray_classic for every pixel (rtc.rs:17-47, tracer.rs:221-259),
    direct light plus ambient at every hit, no photon map; (H, W, 3)."""
    cam = dscene.scene.camera
    out = np.zeros((cam.yreso, cam.xreso, 3))
    amb = np.ascontiguousarray(ambient, np.float64).reshape(3)
    ctx.check(ctx.lib.ppm_trace_classic(ctx.h, dscene.h, int(seed), int(pass_), _dp(amb), _dp(out)))
    return out


def pass_cfg(seed, n_photon, radius0, alpha, max_depth, use_classic=True, pfilter=A.FILTER_NONE, first_pass=1,
             flags=0, pipelines=0):
    """ppm_pass_cfg; pipelines = passes in flight at once (0: automatic)."""
    c = A.PassCfg()
    c.seed, c.n_photon, c.first_pass = int(seed), int(n_photon), int(first_pass)
    c.pipelines = int(pipelines)
    c.radius0, c.alpha = float(radius0), float(alpha)
    c.use_classic, c.max_depth, c.pfilter, c.flags = int(use_classic), int(max_depth), int(pfilter), int(flags)
    return c


class Session:
    """PPM-PA pass loop (util/iterator.rb + util/averager2.rb) on the device."""

    def __init__(self, ctx, dscene, cfg):
        self.ctx, self.dscene, self.cfg = ctx, dscene, cfg
        h = C.c_void_p()
        ctx.check(ctx.lib.ppm_session_create(ctx.h, dscene.h, C.byref(cfg), C.byref(h)))
        self.h = h
        ctx._own(self)

    def run(self, n_pass):
        self.ctx.check(self.ctx.lib.ppm_session_run(self.h, int(n_pass)))

    def image(self, with_var=True):
        cam = self.dscene.scene.camera
        mean = np.zeros((cam.yreso, cam.xreso, 3))
        var = np.zeros_like(mean) if with_var else None
        n = C.c_uint32()
        self.ctx.check(self.ctx.lib.ppm_session_image(self.h, _dp(mean), _dp(var) if with_var else None,
                                                      C.byref(n)))
        return mean, var, n.value

    def radius2(self):
        r = C.c_double()
        self.ctx.check(self.ctx.lib.ppm_session_radius2(self.h, C.byref(r)))
        return r.value

    def set_tile_width(self, width):
        """Eye tiles of width x (64 / width) pixels, width 4, 8, 16 or 32
        (ppm_session_set_tile_width)."""
        self.ctx.check(self.ctx.lib.ppm_session_set_tile_width(self.h, int(width)))

    def reset(self):
        """Back to before the first pass, with the scene's current parameters
        (ppm_session_reset)."""
        self.ctx.check(self.ctx.lib.ppm_session_reset(self.h))

    def backward(self, grad_image):
        """dL/dc1 of every material (n_materials x 3) and dL/d(light_scale)
        for grad_image = dL/d(mean image), by replaying the passes run so far
        (ppm_session_backward)."""
        cam = self.dscene.scene.camera
        g = np.ascontiguousarray(grad_image, np.float64).reshape(cam.yreso, cam.xreso, 3)
        ga = np.zeros((len(self.dscene.scene.materials), 3))
        gl = C.c_double()
        self.ctx.check(self.ctx.lib.ppm_session_backward(self.h, _dp(g), _dp(ga), C.byref(gl)))
        return ga, gl.value

    def timings(self):
        t = A.Timings()
        self.ctx.check(self.ctx.lib.ppm_session_timings(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in A.Timings._fields_ if k != "reserved"}

    def close(self):
        if self.h:
            self.ctx.lib.ppm_session_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_passes(ctx, dscene, cfg, n_pass):
    s = Session(ctx, dscene, cfg)
    try:
        s.run(n_pass)
        mean, var, _ = s.image()
        return mean, var, s.timings()
    finally:
        s.close()


class Batch:
    """Many scenes, each with its own pass schedule, run together: one launch
    per pass stage over every scene still running (ppm_batch_*).  Each
    scene's image equals what its own Session would produce."""

    def __init__(self, ctx, dscenes, cfgs, n_passes):
        self.ctx, self.dscenes = ctx, list(dscenes)
        n = len(self.dscenes)
        hs = (C.c_void_p * n)(*[d.h.value for d in self.dscenes])
        cf = (A.PassCfg * n)(*cfgs)
        npass = (C.c_uint32 * n)(*[int(k) for k in n_passes])
        h = C.c_void_p()
        ctx.check(ctx.lib.ppm_batch_create(ctx.h, n, hs, cf, npass, C.byref(h)))
        self.h = h
        ctx._own(self)
        self.n_passes = [int(k) for k in n_passes]

    def run(self, max_rounds=0):
        self.ctx.check(self.ctx.lib.ppm_batch_run(self.h, int(max_rounds)))

    def launch(self):
        lc = A.LaunchCfg()
        self.ctx.check(self.ctx.lib.ppm_batch_get_launch(self.h, C.byref(lc)))
        return {"eye_waves_per_cu": lc.eye_waves_per_cu, "shade_blocks_per_cu": lc.shade_blocks_per_cu}

    def set_launch(self, eye_waves_per_cu=0, shade_blocks_per_cu=0, tile_width=0):
        lc = A.LaunchCfg()
        lc.eye_waves_per_cu, lc.shade_blocks_per_cu = int(eye_waves_per_cu), int(shade_blocks_per_cu)
        lc.tile_width = int(tile_width)
        self.ctx.check(self.ctx.lib.ppm_batch_set_launch(self.h, C.byref(lc)))

    def launch(self):
        lc = A.LaunchCfg()
        self.ctx.check(self.ctx.lib.ppm_batch_get_launch(self.h, C.byref(lc)))
        return {"eye_waves_per_cu": lc.eye_waves_per_cu, "shade_blocks_per_cu": lc.shade_blocks_per_cu,
                "tile_width": lc.tile_width}

    def tune(self, rounds=2):
        """Autotune the launch sizes and the eye's tile shape on `rounds`
        timed passes per candidate; every scene is reset to before its first
        pass afterwards."""
        lc = A.LaunchCfg()
        self.ctx.check(self.ctx.lib.ppm_batch_tune(self.h, int(rounds), C.byref(lc)))
        return {"eye_waves_per_cu": lc.eye_waves_per_cu, "shade_blocks_per_cu": lc.shade_blocks_per_cu,
                "tile_width": lc.tile_width}

    def image(self, i, with_var=True):
        cam = self.dscenes[i].scene.camera
        mean = np.zeros((cam.yreso, cam.xreso, 3))
        var = np.zeros_like(mean) if with_var else None
        n = C.c_uint32()
        self.ctx.check(self.ctx.lib.ppm_batch_image(self.h, int(i), _dp(mean), _dp(var) if with_var else None,
                                                    C.byref(n)))
        return mean, var, n.value

    def timings(self):
        t = A.Timings()
        self.ctx.check(self.ctx.lib.ppm_batch_timings(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in A.Timings._fields_ if k != "reserved"}

    def close(self):
        if self.h:
            self.ctx.lib.ppm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
