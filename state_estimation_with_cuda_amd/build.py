"""Build libsle.so (SLE_Real = float) and libsle_f64.so (SLE_Real = double).

Host C++ is compiled with g++, the device translation unit with hipcc for
gfx950, and both ABI flavours are linked with hipcc into
state_estimation_with_cuda_amd/lib/.  Usage: ``python -m state_estimation_with_cuda_amd.build``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
SRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CXX = shutil.which("g++") or "g++"
ARCH = os.environ.get("SLE_OFFLOAD_ARCH", "gfx950")

CXXFLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", f"-I{INC}", f"-I{SRC}"]
HIPFLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-result",
    f"-I{INC}",
    f"-I{SRC}",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose: bool = False) -> list[str]:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    jobs = []
    host = ["symbolic.cpp", "engine.cpp"]
    for f in host:
        jobs.append([CXX, *CXXFLAGS, "-c", os.path.join(SRC, f), "-o", os.path.join(OBJ, f + ".o")])
    jobs.append([CXX, *CXXFLAGS, "-c", os.path.join(SRC, "sle_api.cpp"), "-o", os.path.join(OBJ, "api_f32.o")])
    jobs.append(
        [CXX, *CXXFLAGS, "-DSLE_USE_DOUBLE", "-c", os.path.join(SRC, "sle_api.cpp"), "-o", os.path.join(OBJ, "api_f64.o")]
    )
    jobs.append([HIPCC, *HIPFLAGS, "-c", os.path.join(SRC, "device.hip"), "-o", os.path.join(OBJ, "device.o")])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    common = [os.path.join(OBJ, f + ".o") for f in host] + [os.path.join(OBJ, "device.o")]
    libs = []
    for api, name in (("api_f32.o", "libsle.so"), ("api_f64.o", "libsle_f64.so")):
        out = os.path.join(LIB, name)
        tmp = out + ".tmp"
        _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, os.path.join(OBJ, api), *common])
        os.replace(tmp, out)
        libs.append(out)
    return libs


if __name__ == "__main__":
    for p in build(verbose="-v" in sys.argv):
        print(p)
