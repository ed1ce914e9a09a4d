"""Teardown order of the Python binding at interpreter exit.  Objects made on
a context (scenes, photon sets, maps, sessions, batches) destroy themselves
by waiting on the context's stream, so ppm_destroy must come after them.
Left to the interpreter's shutdown, the garbage collector finalized the
module's objects in creation order, context first, and the later destroy
calls read a freed context.  CPU only: the library is replaced by a
recorder of the destroy calls, in a child interpreter that exits with the
objects alive."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r'''
import os, sys
sys.path.insert(0, sys.argv[1])
import ppmpa_amd._abi as A

class Fn:
    def __init__(self, name, w=os.write):
        self.name, self.w = name, w
    def __call__(self, *a):
        if self.name.endswith("destroy"):
            self.w(1, (self.name + "\n").encode())
        return 0

class FakeLib:
    def __getattr__(self, n):
        return Fn(n)

A.load = lambda: FakeLib()
import ctypes as C
import ppmpa_amd as P
byref = C.byref
P.tracer.C.byref = lambda x, *r: (setattr(x, "value", 1234) if isinstance(x, C.c_void_p) else None) or byref(x, *r)
c = P.Context(0)
ds = P.DeviceScene(c, P.two_spheres_scene(8, 8))
ph = P.PhotonSet(c, 10)
m = P.PhotonMap(c, 10)
s = P.Session(c, ds, P.pass_cfg(seed=1, n_photon=10, radius0=0.1, alpha=0.7, max_depth=2))
b = P.Batch(c, [ds], [P.pass_cfg(seed=1, n_photon=10, radius0=0.1, alpha=0.7, max_depth=2)], [2])
'''


def test_context_destroyed_last_at_exit():
    out = subprocess.run([sys.executable, "-c", SCRIPT, ROOT], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    calls = out.stdout.split()
    assert sorted(calls) == sorted(["ppm_batch_destroy", "ppm_session_destroy", "ppm_map_destroy",
                                    "ppm_photons_destroy", "ppm_scene_destroy", "ppm_destroy"]), calls
    assert calls[-1] == "ppm_destroy", calls
    assert calls.index("ppm_batch_destroy") < calls.index("ppm_scene_destroy")
    assert calls.index("ppm_session_destroy") < calls.index("ppm_scene_destroy")


def test_closing_a_context_closes_its_objects_first():
    code = SCRIPT + "\nc.close()\nos.write(1, b'closed\\n')\n"
    out = subprocess.run([sys.executable, "-c", code, ROOT], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    calls = out.stdout.split()
    assert calls[-2:] == ["ppm_destroy", "closed"], calls
    assert len(calls) == 7  # nothing destroyed twice after the explicit close
