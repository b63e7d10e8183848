"""Builds libnsp_amd.so (HIP kernels for gfx950 + the C ABI) in-tree.

    python -m nsp_amd.build            # compile every csrc/*.hip, link the .so

hipcc cross-compiles for gfx950 without a GPU.  Objects go to
nsp_amd/_build/ (not committed); the shared library lands next to this file
so it travels with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnsp_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NSP_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions stay off: hipcc forms them from scalar code by
    # itself and some of their forms have produced wrong results on these
    # boxes; the host pass warns that it does not know the feature.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unknown-warning-option",
    "-I", os.path.join(os.path.dirname(HERE), "include"),
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    deps = [src] + glob.glob(os.path.join(CSRC, "*.hpp")) + \
        glob.glob(os.path.join(os.path.dirname(HERE), "include", "*.h"))
    if os.path.exists(obj) and os.path.getmtime(obj) >= max(os.path.getmtime(d) for d in deps) \
            and not os.environ.get("NSP_FORCE_REBUILD"):
        return obj
    cmd = [HIPCC, *HIP_FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return obj


def build(force: bool = False, jobs: int = 4) -> str:
    os.makedirs(BUILD, exist_ok=True)
    if force:
        os.environ["NSP_FORCE_REBUILD"] = "1"
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libnsp_amd.so failed")
    os.replace(tmp, LIB)
    _build_cpp()
    build_cpp_tests()
    return LIB


CPP_LIB = os.path.join(HERE, "libnsp_amd_cpp.so")
DRIVER = os.path.join(HERE, "nsp_amd_run")


def _build_cpp() -> None:
    """The C++ host API (include/nsp_amd.hpp) as libnsp_amd_cpp.so over
    libnsp_amd.so, and the nsp_amd_run driver; host code, built with hipcc
    for the HIP runtime headers and libraries."""
    inc = os.path.join(os.path.dirname(HERE), "include")
    common = ["-O2", "-std=c++17", "-fPIC", "-I", inc, "-Wno-unused-result"]
    link = ["-L", HERE, "-Wl,-rpath,$ORIGIN"]
    steps = [
        [HIPCC, *common, "-shared", "-o", CPP_LIB + ".tmp", os.path.join(CSRC, "..", "cpp", "api.cpp"),
         *link, "-lnsp_amd"],
        [HIPCC, *common, "-o", DRIVER + ".tmp", os.path.join(CSRC, "..", "cpp", "driver.cpp"),
         *link, "-lnsp_amd_cpp", "-lnsp_amd"],
    ]
    for cmd, out in zip(steps, (CPP_LIB, DRIVER)):
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"build of {os.path.basename(out)} failed")
        os.replace(out + ".tmp", out)


CPP_TESTS = os.path.join(os.path.dirname(HERE), "tests", "cpp")


def build_cpp_tests() -> None:
    """The C++ test programs under tests/cpp (host code written against
    include/nsp_amd.hpp in the reference's spelling), built with -Wall
    -Werror next to their sources and linked to the libraries above."""
    inc = os.path.join(os.path.dirname(HERE), "include")
    for src in sorted(glob.glob(os.path.join(CPP_TESTS, "*.cpp"))):
        out = src[:-4]
        cmd = [HIPCC, "-O2", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", inc, "-o", out + ".tmp", src,
               "-L", HERE, "-Wl,-rpath,$ORIGIN/../../nsp_amd",
               "-lnsp_amd_cpp", "-lnsp_amd"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"build of {os.path.basename(src)} failed")
        os.replace(out + ".tmp", out)


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
