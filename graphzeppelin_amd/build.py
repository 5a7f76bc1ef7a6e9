"""Build libgz_hip.so (the HIP library behind include/gz_hip.h) in-tree for gfx950.

Usage: python -m graphzeppelin_amd.build  (or call build()).
The library lands next to this file, so it travels with the source tree to a GPU host.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgz_hip.so")
SOURCES = ["gz_kernels.hip", "gz_runtime.cpp", "gz_serving.cpp", "gz_noalloc.hip"]
ARCH = os.environ.get("GZ_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: keep the compiler from emitting packed FP32 ALU ops.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(REPO, "include"),
]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _compile(src: str, obj: str) -> None:
    cmd = [_hipcc(), *HIP_FLAGS]
    if not src.endswith(".hip"):
        cmd += ["-x", "hip"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (res.stdout + res.stderr).splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{out}")
    if out.strip():
        sys.stderr.write(out + "\n")


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(objdir, os.path.splitext(s)[0] + ".o") for s in srcs]
    with ThreadPoolExecutor(max_workers=min(4, len(srcs))) as ex:
        list(ex.map(lambda so: _compile(*so), zip(srcs, objs)))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed\n" + res.stdout + res.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_cpp_tests(verbose: bool = False) -> str:
    """Compile tests/cpp/driver_test.cpp (reference-shaped C++ tests of the facade in
    include/graphzeppelin_amd.hpp) against libgz_hip.so."""
    src = os.path.join(REPO, "tests", "cpp", "driver_test.cpp")
    out = os.path.join(REPO, "tests", "cpp", "driver_test")
    cmd = ["g++", "-std=c++14", "-O2", "-Wall", "-I", os.path.join(REPO, "include"), src, "-o", out,
           "-L", HERE, "-lgz_hip", "-Wl,-rpath,$ORIGIN/../../graphzeppelin_amd", "-lpthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("C++ facade test build failed\n" + res.stdout + res.stderr)
    if verbose:
        print("built", out)
    return out


def build_harness(verbose: bool = False) -> str:
    """Compile the end-to-end stream harness (csrc/gz_process_stream.cpp, the counterpart of the
    reference's tools/process_stream.cpp) against libgz_hip.so, next to the library."""
    src = os.path.join(CSRC, "gz_process_stream.cpp")
    out = os.path.join(HERE, "gz_process_stream")
    cmd = ["g++", "-std=c++14", "-O2", "-Wall", "-I", os.path.join(REPO, "include"), src, "-o", out,
           "-L", HERE, "-lgz_hip", "-Wl,-rpath,$ORIGIN", "-lpthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("stream harness build failed\n" + res.stdout + res.stderr)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
    build_cpp_tests(verbose=True)
    build_harness(verbose=True)
