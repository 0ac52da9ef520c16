"""Compile the HIP sources of stormm_amd into stormm_amd/libstormm_amd.so for gfx950."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libstormm_amd.so")
SOURCES = ["mm_kernels.hip", "mm_minimize.hip", "probes.hip", "synthesis.cpp", "plan.cpp", "pool.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("STORMM_AMD_ARCH", "gfx950")

# Packed FP32 instructions are disabled for all device code (-packed-fp32-ops).  Floating-point
# contraction is limited to single source expressions (-ffp-contract=on), so a multiply and an
# add in different statements are never fused: results then follow from the source alone and
# are the same in every kernel that inlines a stage.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-ffp-contract=on"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall",
          "-Wno-unused-result"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + DEVICE_FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    subprocess.run(cmd, check=True)
    return obj


ROOT = os.path.dirname(HERE)
CAPI_DIR = os.path.join(ROOT, "tests", "capi")
CAPI_BIN = os.path.join(CAPI_DIR, "bin")
# Test programs written against the C-ABI (tests/capi/*.cpp): the host check of the 95-bit
# helpers, and a C++ caller that owns its device arrays.
CAPI_PROGRAMS = {"int95_host": [], "kit_boundary": ["-L", HERE, "-lstormm_amd",
                                                    "-Wl,-rpath,$ORIGIN/../../../stormm_amd"]}


def build_capi_programs(verbose: bool = False) -> None:
    os.makedirs(CAPI_BIN, exist_ok=True)
    for name, extra in CAPI_PROGRAMS.items():
        out = os.path.join(CAPI_BIN, name)
        cmd = [HIPCC, "-O2", "-std=c++17", "-x", "hip", f"--offload-arch={ARCH}",
               os.path.join(CAPI_DIR, name + ".cpp"), "-o", out] + extra
        subprocess.run(cmd, check=True)
        if verbose:
            print(f"built {out}")


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_capi_programs(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
