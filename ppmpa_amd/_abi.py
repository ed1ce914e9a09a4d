"""ctypes declarations of include/ppm.h (the C ABI of libppm.so).

The structures mirror the header field for field; tests check that every
function the header declares is exported by the library.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PPMPA_AMD_LIB", os.path.join(HERE, "libppm.so"))

# status codes
PPM_OK = 0
PPM_E_INVALID = -1
PPM_E_NONFINITE = -2
PPM_E_OOM = -3
PPM_E_HIP = -4
PPM_E_RANGE = -5
PPM_E_PARSE = -6
PPM_E_NODEVICE = -7

SHAPE_POINT, SHAPE_PLAIN, SHAPE_SPHERE, SHAPE_POLYGON, SHAPE_PARALLELOGRAM = range(5)
SURF_NOTHING, SURF_SIMPLE, SURF_TS, SURF_DISNEY, SURF_BRADY = range(5)
LIGHT_POINT, LIGHT_PARALLELOGRAM, LIGHT_SUN = range(3)
FILTER_NONE, FILTER_CONE, FILTER_GAUSS = range(3)
WL_RED, WL_GREEN, WL_BLUE = range(3)

SCENE_BVH_AUTO, SCENE_BVH_OFF, SCENE_BVH_ON = range(3)

SESSION_STAGE_TIMING = 1
SESSION_NO_GRAPH = 2

D3 = C.c_double * 3


class ShapeDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("material", C.c_int32),
        ("pos", D3),
        ("nvec", D3),
        ("dir1", D3),
        ("dir2", D3),
        ("dist", C.c_double),
        ("radius", C.c_double),
    ]


class MaterialDesc(C.Structure):
    _fields_ = [
        ("emittance", D3),
        ("transmittance", D3),
        ("ior", D3),
        ("surface", C.c_int32),
        ("reserved", C.c_int32),
        ("color1", D3),
        ("color2", D3),
        ("diffuseness", C.c_double),
        ("metalness", C.c_double),
        ("roughness", C.c_double),
    ]


class LightDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("reserved", C.c_int32),
        ("color", D3),
        ("flux", C.c_double),
        ("pos", D3),
        ("nvec", D3),
        ("dir1", D3),
        ("dir2", D3),
        ("dir", D3),
    ]


class CameraDesc(C.Structure):
    _fields_ = [
        ("xreso", C.c_int32),
        ("yreso", C.c_int32),
        ("progressive", C.c_int32),
        ("antialias", C.c_int32),
        ("blur", C.c_int32),
        ("pfilter", C.c_int32),
        ("eye_pos", D3),
        ("origin", D3),
        ("esx", D3),
        ("esy", D3),
        ("eex", D3),
        ("eey", D3),
    ]


class SceneDesc(C.Structure):
    _fields_ = [
        ("n_shapes", C.c_int32),
        ("n_materials", C.c_int32),
        ("n_lights", C.c_int32),
        ("flags", C.c_int32),
        ("shapes", C.POINTER(ShapeDesc)),
        ("materials", C.POINTER(MaterialDesc)),
        ("lights", C.POINTER(LightDesc)),
        ("camera", CameraDesc),
        ("bounds_min", D3),
        ("bounds_max", D3),
    ]


class PassCfg(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("n_photon", C.c_uint64),
        ("first_pass", C.c_uint32),
        ("pipelines", C.c_uint32),
        ("radius0", C.c_double),
        ("alpha", C.c_double),
        ("use_classic", C.c_int32),
        ("max_depth", C.c_int32),
        ("pfilter", C.c_int32),
        ("flags", C.c_int32),
    ]


class Timings(C.Structure):
    _fields_ = [
        ("ms_photon_trace", C.c_double),
        ("ms_map_build", C.c_double),
        ("ms_eye_trace", C.c_double),
        ("ms_gather", C.c_double),
        ("ms_accumulate", C.c_double),
        ("ms_total", C.c_double),
        ("photons_emitted", C.c_uint64),
        ("photons_stored", C.c_uint64),
        ("gather_queries", C.c_uint64),
        ("passes", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class LaunchCfg(C.Structure):
    _fields_ = [
        ("eye_waves_per_cu", C.c_uint32),
        ("shade_blocks_per_cu", C.c_uint32),
        ("tile_width", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


vp = C.c_void_p
dp = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
st = C.c_int32

# name -> (restype, argtypes)
PROTOTYPES = {
    "ppm_create": (st, [C.POINTER(vp), C.c_int, vp]),
    "ppm_destroy": (None, [vp]),
    "ppm_last_error": (C.c_char_p, [vp]),
    "ppm_status_string": (C.c_char_p, [st]),
    "ppm_sync": (st, [vp]),
    "ppm_camera_setup": (st, [C.POINTER(CameraDesc), C.c_int32, C.c_int32, dp, dp, dp,
                              C.c_double, C.c_double, C.c_double]),
    "ppm_scene_create": (st, [vp, C.POINTER(SceneDesc), C.POINTER(vp)]),
    "ppm_scene_destroy": (None, [vp]),
    "ppm_scene_stats": (st, [vp, u32p, u32p, u32p]),
    "ppm_intersect": (st, [vp, vp, dp, dp, C.c_uint64, dp, C.POINTER(C.c_int32)]),
    "ppm_photons_create": (st, [vp, C.c_uint64, C.POINTER(vp)]),
    "ppm_photons_destroy": (None, [vp]),
    "ppm_trace_photons": (st, [vp, vp, C.c_uint64, C.c_uint32, C.c_uint64, C.c_int32, C.c_int32, vp]),
    "ppm_photons_count": (st, [vp, vp, u64p]),
    "ppm_photons_upload": (st, [vp, vp, C.c_uint64, dp, dp, u8p]),
    "ppm_photons_download": (st, [vp, vp, C.c_uint64, dp, dp, u8p, u32p, u64p]),
    "ppm_map_create": (st, [vp, C.c_uint64, C.POINTER(vp)]),
    "ppm_map_destroy": (None, [vp]),
    "ppm_map_build": (st, [vp, vp, vp, C.c_double, vp]),
    "ppm_map_build_grid": (st, [vp, vp, vp, C.c_double, C.c_double, vp]),
    "ppm_radius_query": (st, [vp, vp, dp, C.c_uint64, C.c_double, C.c_int32, C.c_uint64, u32p, u64p]),
    "ppm_gather": (st, [vp, vp, dp, dp, C.c_uint64, C.c_double, C.c_int32, C.c_int32, dp]),
    "ppm_gather_noalloc": (st, [vp, vp, dp, dp, C.c_uint64, C.c_double, C.c_int32, C.c_int32, dp]),
    "ppm_trace_classic": (st, [vp, vp, C.c_uint64, C.c_uint32, dp, dp]),
    "ppm_trace_eye": (st, [vp, vp, vp, C.c_uint64, C.c_uint32, C.c_double, C.c_int32, C.c_int32,
                           C.c_int32, dp]),
    "ppm_session_create": (st, [vp, vp, C.POINTER(PassCfg), C.POINTER(vp)]),
    "ppm_session_destroy": (None, [vp]),
    "ppm_session_run": (st, [vp, C.c_uint32]),
    "ppm_session_image": (st, [vp, dp, dp, u32p]),
    "ppm_session_radius2": (st, [vp, dp]),
    "ppm_session_timings": (st, [vp, C.POINTER(Timings)]),
    "ppm_session_set_tile_width": (st, [vp, C.c_uint32]),
    "ppm_run_passes": (st, [vp, vp, C.POINTER(PassCfg), C.c_uint32, dp, dp, C.POINTER(Timings)]),
    "ppm_scene_set_params": (st, [vp, vp, C.c_uint32, dp, dp, C.c_double]),
    "ppm_session_backward": (st, [vp, dp, dp, dp]),
    "ppm_session_reset": (st, [vp]),
    "ppm_batch_create": (st, [vp, C.c_uint32, C.POINTER(vp), C.POINTER(PassCfg), u32p, C.POINTER(vp)]),
    "ppm_batch_destroy": (None, [vp]),
    "ppm_batch_run": (st, [vp, C.c_uint32]),
    "ppm_batch_set_launch": (st, [vp, C.POINTER(LaunchCfg)]),
    "ppm_batch_get_launch": (st, [vp, C.POINTER(LaunchCfg)]),
    "ppm_batch_tune": (st, [vp, C.c_uint32, C.POINTER(LaunchCfg)]),
    "ppm_batch_image": (st, [vp, C.c_uint32, dp, dp, u32p]),
    "ppm_batch_timings": (st, [vp, C.POINTER(Timings)]),
}

_lib = None


def load(path=LIB_PATH):
    """Load libppm.so; raises when the HIP library has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            "libppm.so is not built (run `python -m ppmpa_amd.build`); the GPU path has no CPU fallback")
    lib = C.CDLL(path)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def header_symbols(header_path=None):
    """Function names declared in include/ppm.h."""
    import re
    if header_path is None:
        header_path = os.path.join(HERE, "..", "include", "ppm.h")
    text = open(header_path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ppm_[a-z0-9_]+)\s*\(", text)))
