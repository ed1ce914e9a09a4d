"""The C ABI library: it loads, exports every function include/ppm.h
declares, its structures match the header's layout, and its host-only entry
points behave.  No GPU compute (this runs on CPU-only machines)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import ppmpa_amd as P
from ppmpa_amd import _abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(A.LIB_PATH):
        pytest.fail("libppm.so not built (python -m ppmpa_amd.build)")
    return A.load()


def test_exports_every_header_symbol(lib):
    names = A.header_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python prototypes cover exactly the header
    assert sorted(A.PROTOTYPES) == names


def test_struct_layout_matches_header(tmp_path):
    """Compile a C probe against include/ppm.h and compare sizes/offsets."""
    src = tmp_path / "probe.c"
    structs = {"ppm_shape_desc": A.ShapeDesc, "ppm_material_desc": A.MaterialDesc, "ppm_light_desc": A.LightDesc,
               "ppm_camera_desc": A.CameraDesc, "ppm_scene_desc": A.SceneDesc, "ppm_pass_cfg": A.PassCfg,
               "ppm_timings": A.Timings}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ppm.h"', "int main(void){"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, stdout=subprocess.PIPE, text=True).stdout.split("\n")
    got = dict(l.split() for l in out if l)
    for cname, py in structs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(py, fname).offset, (cname, fname)


def test_camera_setup_matches_read_camera(lib):
    """ppm_camera_setup (C) == Camera.screen() (camera.rs:151-164)."""
    cam = P.read_camera()
    d = A.CameraDesc()
    eye, tgt, up = (C.c_double * 3)(*cam.eye), (C.c_double * 3)(*cam.target), (C.c_double * 3)(*cam.upper)
    assert lib.ppm_camera_setup(C.byref(d), cam.xreso, cam.yreso, eye, tgt, up, cam.focus, cam.focal_len,
                                cam.f_number) == 0
    ref = cam.screen()
    for k in ("origin", "esx", "esy", "eex", "eey"):
        assert np.allclose(list(getattr(d, k)), ref[k], rtol=1e-14, atol=1e-16), k
    # invalid arguments are rejected, not aborted on
    assert lib.ppm_camera_setup(C.byref(d), 0, 10, eye, tgt, up, 1.0, 0.05, 4.0) == A.PPM_E_INVALID
    assert lib.ppm_camera_setup(C.byref(d), 10, 10, eye, eye, up, 1.0, 0.05, 4.0) == A.PPM_E_INVALID


def test_status_strings_and_null_handles(lib):
    for s in range(-7, 1):
        assert lib.ppm_status_string(s)
    assert lib.ppm_last_error(None) == b"null context"
    assert lib.ppm_sync(None) == A.PPM_E_INVALID
    assert lib.ppm_session_run(None, 1) == A.PPM_E_INVALID
    lib.ppm_destroy(None)
    lib.ppm_scene_destroy(None)
    lib.ppm_photons_destroy(None)
    lib.ppm_map_destroy(None)
    lib.ppm_session_destroy(None)


def test_create_without_device_fails_cleanly(lib):
    """On a machine without a gfx950 device the context refuses to open."""
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK):
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert lib.ppm_create(C.byref(h), 0, None) == A.PPM_E_NODEVICE
    assert not h.value
