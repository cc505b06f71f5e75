"""Compile the HIP engine and its C ABI into vmps_amd/libvmps_hxv.so (gfx950).

Run as ``python -m vmps_amd.build``.  Every HIP translation unit is compiled
with packed FP32 instructions disabled; objects go to vmps_amd/_build/.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libvmps_hxv.so")
BUILD = os.path.join(HERE, "_build")
ARCH = os.environ.get("VMPS_OFFLOAD_ARCH", "gfx950")

SOURCES = ["sym.cpp", "pivot.cpp", "env.cpp", "capi.cpp", "capi_env.cpp", "split.cpp", "engine.hip",
           "engine_bf16.hip", "lanczos.hip", "lanczos_batch.hip", "svd.hip"]
# C++ units that call the HIP runtime API are compiled in HIP mode
HIP_HOST_UNITS = {"capi.cpp", "capi_env.cpp", "split.cpp"}

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]
HIP_FLAGS = [
    "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".hip") or src in HIP_HOST_UNITS:
        cmd = [_hipcc(), "-c", "-x", "hip", path, "-o", obj] + COMMON + HIP_FLAGS
    else:
        # host-only translation units: plain C++ through hipcc's host compiler
        cmd = [_hipcc(), "-c", "-x", "c++", path, "-o", obj] + COMMON
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), res.stdout, res.stderr))
    noise = [l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    # rocBLAS / rocSOLVER: the dense GEMM and eigensolver of the device split (split.cpp)
    cmd = [_hipcc(), "-shared", "-o", tmp] + objs + ["--offload-arch=" + ARCH, "-lamdhip64", "-lrocsolver",
                                                     "-lrocblas"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), res.stdout, res.stderr))
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build()
