"""Build recipe of the native library feelpp_amd/libfeelpp_hip.so (gfx950).

Compiles the HIP device layer (csrc/device/*.hip) with hipcc for gfx950 and
the host side (csrc/host/*.cpp) with g++, then links one shared library with
both C-ABIs (include/feelpp_hip.h, include/feelpp_host.h).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfeelpp_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FEELPP_AMD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
HIP_FLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]
CXX_FLAGS = ["-march=x86-64-v3", "-Wall", "-Wno-unused-parameter"]


def _sources():
    dev = sorted(os.path.join(CSRC, "device", f) for f in os.listdir(os.path.join(CSRC, "device")) if f.endswith(".hip"))
    host = sorted(os.path.join(CSRC, "host", f) for f in os.listdir(os.path.join(CSRC, "host")) if f.endswith(".cpp"))
    return dev, host


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip", "-c", src, "-o", obj] + COMMON + HIP_FLAGS
    else:
        cmd = ["g++", "-c", src, "-o", obj] + COMMON + CXX_FLAGS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    dev, host = _sources()
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, dev + host))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-o", tmp] + objs + ["--offload-arch=" + ARCH, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    # quickstart driver executables, linked against the library (rpath to feelpp_amd/)
    bindir = os.path.join(HERE, "bin")
    os.makedirs(bindir, exist_ok=True)
    apps = sorted(f for f in os.listdir(os.path.join(HERE, "apps")) if f.endswith(".cpp"))
    for app in apps:
        exe = os.path.join(bindir, "feelpp_" + app[:-4])
        cmd = ["g++", os.path.join(HERE, "apps", app), "-o", exe + ".tmp"] + COMMON + CXX_FLAGS + \
              [LIB, "-Wl,-rpath,$ORIGIN/..", "-lpthread"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("app build failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
        os.replace(exe + ".tmp", exe)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
