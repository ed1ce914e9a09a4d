"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, always as the checker.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from ppmpa_amd import _abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.environ.get("PPM_ORACLE_SO") or os.path.join(ORACLE_DIR, "liboracle.so")

dp = C.POINTER(C.c_double)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
u8p = C.POINTER(C.c_uint8)
vp = C.c_void_p

_PROTOS = {
    "orc_philox": (None, [u32p, u32p, u32p]),
    "orc_filter_cone": (C.c_double, [C.c_double, C.c_double]),
    "orc_filter_gauss": (C.c_double, [C.c_double, C.c_double]),
    "orc_normalize": (C.c_int, [dp, dp]),
    "orc_ray_target": (C.c_int, [dp, dp, C.c_double, dp]),
    "orc_polygon_normal": (C.c_int, [dp, dp, dp, dp]),
    "orc_color_normalize": (None, [dp, dp]),
    "orc_decide_wavelength": (C.c_int, [dp, C.c_double]),
    "orc_check_under": (C.c_int, [dp, C.c_int, C.c_double]),
    "orc_schlick": (C.c_double, [C.c_double, C.c_double]),
    "orc_relative_ior": (C.c_double, [C.c_double, C.c_double]),
    "orc_specular_refraction": (C.c_int, [dp, dp, C.c_double, dp, dp]),
    "orc_uniforms": (None, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, dp]),
    "orc_diffuse_samples": (None, [dp, C.c_uint64, C.c_int, dp]),
    "orc_scene_create": (vp, [C.POINTER(A.SceneDesc)]),
    "orc_scene_destroy": (None, [vp]),
    "orc_intersect": (C.c_int, [vp, dp, dp, dp, dp]),
    "orc_direct_light": (None, [vp, dp, dp, dp]),
    "orc_trace_photons": (C.c_int64, [vp, C.c_uint64, C.c_uint32, C.c_uint64, C.c_int, C.c_int, C.c_uint64, dp,
                                      dp, u8p, u32p, dp]),
    "orc_within_brute": (C.c_int64, [C.c_uint64, dp, C.c_uint64, dp, C.c_double, C.c_int, C.c_uint64, u32p, u64p]),
    "orc_map_build": (vp, [C.c_uint64, dp, dp, u8p, C.c_double]),
    "orc_map_destroy": (None, [vp]),
    "orc_map_within": (C.c_int64, [vp, dp, C.c_double, C.c_uint64, u32p, dp]),
    "orc_gather": (None, [vp, C.c_uint64, dp, dp, C.c_double, C.c_int, dp]),
    "orc_gather_f32set": (None, [vp, C.c_uint64, dp, dp, C.c_double, C.c_int, dp]),
    "orc_trace_eye": (None, [vp, vp, C.c_uint64, C.c_uint32, C.c_double, C.c_int, C.c_int, C.c_int, dp]),
    "orc_trace_classic": (None, [vp, C.c_uint64, C.c_uint32, dp, dp]),
    "orc_radius_schedule": (None, [C.c_double, C.c_double, C.c_uint32, C.c_uint32, dp]),
    "orc_run_passes": (C.c_int64, [vp, C.POINTER(A.PassCfg), C.c_uint32, dp, dp, C.c_int, dp, dp]),
    "orc_max_threads": (C.c_int, []),
    "orc_reflection_glossy": (None, [dp, dp, C.c_double, C.c_uint64, C.c_uint32, dp]),
    "orc_next_direction": (None, [C.POINTER(A.MaterialDesc), C.c_double, dp, dp, C.c_int, C.c_uint64, C.c_uint32,
                                  C.POINTER(C.c_int32), dp]),
    "orc_scene_set_params": (None, [vp, C.c_int, dp, dp, C.c_double]),
    "orc_run_passes_grad": (C.c_int64, [vp, C.POINTER(A.PassCfg), C.c_uint32, dp, dp, dp, C.c_int]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-C", ORACLE_DIR], check=True, stdout=subprocess.PIPE)
        L = C.CDLL(ORACLE_SO)
        for k, (r, a) in _PROTOS.items():
            f = getattr(L, k)
            f.restype, f.argtypes = r, a
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(dp)


def arr3(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox(c, k, o)
    return list(o)


class OScene:
    def __init__(self, scene):
        desc, self._keep = scene.desc()
        self.h = lib().orc_scene_create(C.byref(desc))
        self.scene = scene

    def __del__(self):
        try:
            lib().orc_scene_destroy(self.h)
        except Exception:
            pass

    def intersect(self, o, d):
        pos, nrm = np.zeros(3), np.zeros(3)
        obj = lib().orc_intersect(self.h, _d(arr3(o)), _d(arr3(d)), _d(pos), _d(nrm))
        return obj, pos, nrm

    def direct_light(self, p, n):
        out = np.zeros(3)
        lib().orc_direct_light(self.h, _d(arr3(p)), _d(arr3(n)), _d(out))
        return out

    def trace_photons(self, seed, pass_, n_photon, uc=True, maxd=10):
        cap = int(n_photon) * int(maxd) + 16
        pos, dir = np.zeros((cap, 3)), np.zeros((cap, 3))
        wl, meta = np.zeros(cap, np.uint8), np.zeros(cap, np.uint32)
        pw = C.c_double()
        n = lib().orc_trace_photons(self.h, seed, pass_, n_photon, int(uc), maxd, cap, _d(pos), _d(dir),
                                    wl.ctypes.data_as(u8p), meta.ctypes.data_as(u32p), C.byref(pw))
        assert n >= 0
        return pos[:n], dir[:n], wl[:n], meta[:n], pw.value

    def trace_eye(self, omap, seed, pass_, r2, uc=True, maxd=10, pf=0):
        cam = self.scene.camera
        out = np.zeros((cam.yreso, cam.xreso, 3))
        lib().orc_trace_eye(self.h, omap.h, seed, pass_, r2, int(uc), maxd, pf, _d(out))
        return out

    def trace_classic(self, seed, pass_, ambient=(0.0, 0.0, 0.0)):
        """rtc: trace_ray_classic for every pixel (H, W, 3)."""
        cam = self.scene.camera
        out = np.zeros((cam.yreso, cam.xreso, 3))
        amb = np.ascontiguousarray(ambient, np.float64)
        lib().orc_trace_classic(self.h, seed, pass_, _d(amb), _d(out))
        return out

    def set_params(self, c1=None, q=None, light_scale=1.0):
        """c1 (n_mat x 3) and survival probabilities q of the diffuse bounce
        (None: q = c1); light_scale multiplies every flux and emittance."""
        n = len(self.scene.materials)
        c1a = None if c1 is None else np.ascontiguousarray(c1, np.float64).reshape(n, 3)
        qa = None if q is None else np.ascontiguousarray(q, np.float64).reshape(n, 3)
        lib().orc_scene_set_params(self.h, n, _d(c1a) if c1a is not None else None,
                                   _d(qa) if qa is not None else None, float(light_scale))

    def run_passes_grad(self, cfg, n_pass, g_img, threads=0):
        """dL/dc1 (n_mat x 3) of the passes' estimator for dL/d(mean image)
        g_img, and the mean image."""
        cam = self.scene.camera
        g = np.ascontiguousarray(g_img, np.float64).reshape(cam.yreso, cam.xreso, 3)
        grad = np.zeros((len(self.scene.materials), 3))
        mean = np.zeros((cam.yreso, cam.xreso, 3))
        lib().orc_run_passes_grad(self.h, C.byref(cfg), n_pass, _d(g), _d(grad), _d(mean), threads)
        return grad, mean

    def run_passes(self, cfg, n_pass, threads=0):
        cam = self.scene.camera
        mean = np.zeros((cam.yreso, cam.xreso, 3))
        var = np.zeros_like(mean)
        t1, t2 = C.c_double(), C.c_double()
        stored = lib().orc_run_passes(self.h, C.byref(cfg), n_pass, _d(mean), _d(var), threads, C.byref(t1),
                                      C.byref(t2))
        return mean, var, stored, t1.value, t2.value


class OMap:
    def __init__(self, pos, dir, wl, power):
        self.pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        self.dir = np.ascontiguousarray(dir, dtype=np.float64).reshape(-1, 3)
        self.wl = np.ascontiguousarray(wl, dtype=np.uint8)
        self.power = float(power)
        self.h = lib().orc_map_build(self.pos.shape[0], _d(self.pos), _d(self.dir), self.wl.ctypes.data_as(u8p),
                                     float(power))

    def __del__(self):
        try:
            lib().orc_map_destroy(self.h)
        except Exception:
            pass

    def within(self, q, r2):
        cap = self.pos.shape[0] + 1
        idx = np.zeros(cap, np.uint32)
        d2 = np.zeros(cap)
        n = lib().orc_map_within(self.h, _d(arr3(q)), float(r2), cap, idx.ctypes.data_as(u32p), _d(d2))
        return idx[:n], d2[:n]

    def gather(self, q, nrm, r2, pf=0, f32set=False):
        """estimate_radiance in FP64 over the FP64 ball, or (f32set) over the
        ball that the GPU passes' float predicate decides."""
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
        nrm = np.ascontiguousarray(nrm, dtype=np.float64).reshape(-1, 3)
        out = np.zeros_like(q)
        f = lib().orc_gather_f32set if f32set else lib().orc_gather
        f(self.h, q.shape[0], _d(q), _d(nrm), float(r2), int(pf), _d(out))
        return out


def within_brute(pos, q, r2, mode=0):
    """mode 0: f64 (the reference's predicate), 1: f32 as on the GPU."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
    nq = q.shape[0]
    off = np.zeros(nq + 1, np.uint64)
    cap = 1
    idx = np.zeros(cap, np.uint32)
    n = lib().orc_within_brute(pos.shape[0], _d(pos), nq, _d(q), float(r2), mode, cap, idx.ctypes.data_as(u32p),
                               off.ctypes.data_as(u64p))
    if n < 0:
        cap = -n
        idx = np.zeros(cap, np.uint32)
        n = lib().orc_within_brute(pos.shape[0], _d(pos), nq, _d(q), float(r2), mode, cap, idx.ctypes.data_as(u32p),
                                   off.ctypes.data_as(u64p))
    return [idx[int(off[i]):int(off[i + 1])] for i in range(nq)]


def reflection_glossy(n, r, pw, seed, count):
    """count samples of reflection_glossy (physics.rs:232-261) about r."""
    out = np.zeros((count, 3))
    lib().orc_reflection_glossy(_d(arr3(n)), _d(arr3(r)), float(pw), int(seed), int(count), _d(out))
    return out


def next_direction(material, eta, n, v, wl, seed, count):
    """count draws of Surface::next_direction (surface.rs:211-258): (kinds,
    dirs), kind 0 absorbed, 1 reflected, 2 transmitted."""
    m = material
    d = A.MaterialDesc()
    d.emittance, d.transmittance, d.ior = A.D3(*m.emittance), A.D3(*m.transmittance), A.D3(*m.ior)
    d.surface = int(m.surface)
    d.color1, d.color2 = A.D3(*m.color1), A.D3(*m.color2)
    d.diffuseness, d.metalness, d.roughness = float(m.diffuseness), float(m.metalness), float(m.roughness)
    kinds = np.zeros(count, np.int32)
    dirs = np.zeros((count, 3))
    lib().orc_next_direction(C.byref(d), float(eta), _d(arr3(n)), _d(arr3(v)), int(wl), int(seed), int(count),
                             kinds.ctypes.data_as(C.POINTER(C.c_int32)), _d(dirs))
    return kinds, dirs


def radius_schedule(r1, alpha, first, n):
    out = np.zeros(n)
    lib().orc_radius_schedule(r1, alpha, first, n, _d(out))
    return out
