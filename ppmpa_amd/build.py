"""Build libppm.so (the gfx950 HIP kernels + the C ABI of include/ppm.h) in-tree.

Usage: python -m ppmpa_amd.build   (or build() from Python)
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libppm.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are kept out of the generated code (the host pass
# warns that it does not know the feature; that is harmless).
COMMON = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(HERE, "..", "include"),
]

SOURCES = ["ppm_kernels.hip", "ppm_api.cpp"]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError("command failed: " + " ".join(cmd))
    return proc.stdout


def build(verbose=False, out=LIB, defines=()):
    """Compile every source to an object and link libppm.so next to this file
    (or at `out`, with extra -D `defines`)."""
    if not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found at " + HIPCC)
    objdir = os.path.join(HERE, "_build", os.path.basename(out))
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src + ".o")
        flags = COMMON + ["-D" + d for d in defines]
        cmd = [HIPCC] + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC, "-x", "hip"] + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
        log = _run(cmd)
        if verbose and log.strip():
            sys.stderr.write(log)
        objs.append(obj)
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    shutil.move(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
