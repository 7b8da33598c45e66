"""Build recipe for libmkde.so (HIP kernels for gfx950 + the C-ABI host library)
and the mgpukde_tiled command-line program.

Run as ``python -m mgpukde_tiled_amd.build`` or through ``__graft_entry__.build()``.
Outputs land in the package directory (in-tree), so the built library travels
with the source tree.  Every build recompiles.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmkde.so")
CLI = os.path.join(HERE, "mgpukde_tiled")
RATES = os.path.join(HERE, "mkde_rates")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled for all device code (hipcc-only flag).
DEVICE_FLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-I" + INCLUDE, "-I" + CSRC]
# Host-only translation units use the HIP runtime API through its C++ headers.
HOST_FLAGS = ["-D__HIP_PLATFORM_AMD__=1", "-I/opt/rocm/include"]

HIP_SOURCES = ["mkde_kernels.hip", "mkde_bins.hip"]
HOST_SOURCES = ["mkde_capi.cpp", "mkde_tiff.cpp", "mkde_serve.cpp"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run([HIPCC] + DEVICE_FLAGS + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HOST_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, src + ".o")
        _run([HIPCC] + COMMON + HOST_FLAGS + ["-x", "c++", "-c", path, "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC] + DEVICE_FLAGS + ["-shared", "-fPIC", "-o", tmp] + objs + ["-lz", "-pthread"])
    os.replace(tmp, LIB)
    cli_src = os.path.join(CSRC, "mkde_cli.cpp")
    if os.path.exists(cli_src):
        _run([HIPCC] + COMMON + HOST_FLAGS + ["-x", "c++", cli_src, "-o", CLI, "-L" + HERE, "-lmkde",
                                 "-Wl,-rpath,$ORIGIN"])
    # instruction-rate probe (exp / FMA / MFMA throughput) used by bench.py's roofline
    rates_src = os.path.join(CSRC, "mkde_rates.hip")
    if os.path.exists(rates_src):
        _run([HIPCC] + DEVICE_FLAGS + COMMON + [rates_src, "-o", RATES])
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
