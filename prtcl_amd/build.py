"""Build the HIP library prtcl_amd/_lib/libprtcl_hip.so for gfx950.

Every source compiles with hipcc for --offload-arch=gfx950 only, with packed
FP32 instructions disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
The procedure kernels of sph.hip and wkbb18.hip are built without FMA
contraction so their pair terms round exactly like the FP32 CPU code; the
fused step kernels contract within single expressions only, so every
schedule of a pass rounds alike.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(LIBDIR, "libprtcl_hip.so")

ARCH = "gfx950"
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unknown-pragmas",
]
# sources and their extra flags
SOURCES = {
    "nsearch.hip": [],
    "sph.hip": ["-ffp-contract=off"],
    # contraction only within one expression: the LDS and global-memory
    # schedules of a pass share the pair arithmetic and must round alike
    "fused.hip": ["-ffp-contract=on"],
    "wkbb18.hip": ["-ffp-contract=off"],
    "pt16.hip": ["-ffp-contract=off"],
    "scene_impl.hip": [],
    "serve.hip": [],
    "checkpoint.hip": [],
    "vtk.hip": [],
    "trace.hip": ["-ffp-contract=off"],
    "capi.hip": [],
}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str, extra, verbose: bool) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False, jobs: int = 6) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, SOURCES[s], verbose), srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    build_compat(verbose)
    return LIB


COMPAT_SRC = os.path.join(HERE, "compat", "prtcl_compat.cpp")
COMPAT_LIB = os.path.join(LIBDIR, "libprtcl_compat.so")


def build_compat(verbose: bool = False) -> str:
    """The C++ host interface (Model, Neighborhood, SchemeBase, ... and the
    GPU schemes) as libprtcl_compat.so over libprtcl_hip.so; host-only code,
    built with g++."""
    cxx = os.environ.get("CXX", "g++")
    tmp = COMPAT_LIB + ".tmp"
    cmd = [cxx, "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-Wextra", COMPAT_SRC,
           "-L" + LIBDIR, "-lprtcl_hip", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compat build failed:\n{r.stderr}")
    os.replace(tmp, COMPAT_LIB)
    return COMPAT_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
