"""Builds libmidsx_amd.so (HIP kernels for gfx950 + host table builder +
C-ABI) in-tree under midsx_amd/_lib.

    python -m midsx_amd.build      # or midsx_amd.build.build()
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "_lib")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("MIDSX_AMD_ARCH", "gfx950")

HIPCC = os.path.join(ROCM, "bin", "hipcc")
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-munsafe-fp-atomics",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include")]

SOURCES = [
    ("transport.hip", "hip"),
    ("tables.cpp", "host"),
    ("scene.cpp", "host"),
    ("capi.cpp", "host"),
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed: %s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)

    def one(item):
        src, kind = item
        out = os.path.join(OBJ, src + ".o")
        if kind == "hip":
            cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", out]
        else:
            cmd = ["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, src), "-o", out]
        log = _run(cmd)
        if verbose and log.strip():
            print(log)
        return out

    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(one, SOURCES))
    so = os.path.join(LIB, "libmidsx_amd.so")
    tmp = so + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lz", "-lpthread"])
    os.replace(tmp, so)
    return so


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
