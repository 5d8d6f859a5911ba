"""Memory-safety runs of the GPU tests on the CPU.

The library's own HIP sources (prtcl_amd/csrc/*.hip, unchanged) are compiled
as host C++ against tests/emu/hip/hip_runtime.h, a CPU emulation of the HIP
subset they use, with AddressSanitizer, into a copy of the repository under
tests/emu/_tree whose prtcl_amd/_lib/libprtcl_hip.so is that build; then the
`gpu` tests run there. This is a checker, not a fallback: the library itself
never loads it, and nothing here runs on a GPU.

What the emulation checks (tests/emu/emu_rt.cpp):
* every GPU thread is a fiber; `__syncthreads` waits for the live threads of
  the block, every shuffle / ballot for the live lanes of its 64-lane wave;
  divergent wave operations, threads at different barriers, deadlocks and
  shuffles that read an exited lane are reported;
* device and page-locked allocations are ASan-tracked host memory (hipMalloc
  memory filled with 0xFF, so reads of unwritten memory show), every copy and
  memset must lie inside the allocation it names, and every 8-byte word of a
  kernel argument that looks like a host address must lie inside a device,
  page-locked or symbol allocation;
* captured graphs replay their recorded launches;
* EMU_DEFER=1 runs stream work only at synchronisation points (stream or
  event synchronisation, hipStreamQuery, hipFree), so host code that reuses or
  frees memory a queued copy or kernel still needs, or destroys a graph with
  launches queued, is caught.

Usage:
    python tests/emu/run.py [--defer] [--workers N] [test paths] [pytest args ...]
e.g. python tests/emu/run.py --defer --workers 6 --ignore=tests/test_gpu_configs.py
     python tests/emu/run.py --defer tests/test_gpu_batch.py::test_batch_equals_alone
Reports go to stderr and to tests/emu/_tree/emu_reports.log; ASan reports to
tests/emu/_tree/asan.*. The emulation runs a kernel at roughly 10^5 - 10^6
thread-instructions per second per core: the small-scene tests take minutes,
the full-size configuration tests hours.
"""

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
TREE = os.path.join(HERE, "_tree")
OBJ = os.path.join(HERE, "_obj")
CXX = "/opt/rocm/llvm/bin/clang++"


def asan_runtime() -> str:
    r = subprocess.run([CXX, "-print-file-name=libclang_rt.asan-x86_64.so"],
                       capture_output=True, text=True, check=True)
    return r.stdout.strip()


def build(san: str = "address") -> str:
    sys.path.insert(0, ROOT)
    from prtcl_amd import build as B

    flags = ["-x", "c++", "-std=c++17", "-O2", "-g", "-fno-omit-frame-pointer", "-fPIC",
             "-I" + HERE, "-I" + os.path.join(ROOT, "include"), "-Wno-unused-result",
             "-Wno-unknown-pragmas", "-Wno-pass-failed"]
    if san:
        flags.append("-fsanitize=" + san)
    os.makedirs(OBJ, exist_ok=True)
    if os.path.exists(TREE):
        shutil.rmtree(TREE)
    ign = shutil.ignore_patterns(".git", "_work", "__pycache__", "_build", "*.so",
                                 "emu")
    for d in ("prtcl_amd", "tests", "oracle", "include"):
        shutil.copytree(os.path.join(ROOT, d), os.path.join(TREE, d), ignore=ign)
    for f in os.listdir(ROOT):
        p = os.path.join(ROOT, f)
        if os.path.isfile(p) and f.endswith((".py", ".ini", ".cfg", ".toml")):
            shutil.copy(p, TREE)
    # the oracle is native host code: build it in the copy
    subprocess.check_call(["make", "-s", "-C", os.path.join(TREE, "oracle")])

    def cc(src, extra):
        obj = os.path.join(OBJ, os.path.splitext(os.path.basename(src))[0] + ".o")
        r = subprocess.run([CXX, *flags, *extra, "-c", src, "-o", obj], capture_output=True,
                           text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{src}:\n{r.stderr[-6000:]}")
        return obj

    jobs = [(os.path.join(ROOT, "prtcl_amd", "csrc", s),
             [f for f in B.SOURCES[s] if f.startswith("-ffp")]) for s in B.SOURCES]
    jobs.append((os.path.join(HERE, "emu_rt.cpp"), []))
    with cf.ThreadPoolExecutor(6) as ex:
        objs = list(ex.map(lambda j: cc(*j), jobs))
    lib = os.path.join(TREE, "prtcl_amd", "_lib")
    os.makedirs(lib, exist_ok=True)
    rt = asan_runtime() if san else None
    link = [CXX, "-shared", "-fPIC", *objs, "-o", os.path.join(lib, "libprtcl_hip.so")]
    if rt:
        link += [rt, "-Wl,-rpath," + os.path.dirname(rt)]
    subprocess.check_call(link)
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-g", "-fPIC", "-shared",
                           os.path.join(ROOT, "prtcl_amd", "compat", "prtcl_compat.cpp"),
                           "-L" + lib, "-lprtcl_hip", "-Wl,-rpath,$ORIGIN",
                           "-o", os.path.join(lib, "libprtcl_compat.so")])
    return lib


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--defer", action="store_true", help="run stream work at sync points only")
    ap.add_argument("--workers", type=int, default=1)
    ap.add_argument("--no-build", action="store_true")
    args, rest = ap.parse_known_args()
    if not args.no_build:
        build()
    env = dict(os.environ)
    env["LD_PRELOAD"] = asan_runtime()
    env["ASAN_OPTIONS"] = ("detect_leaks=0:abort_on_error=1:halt_on_error=1:log_path=" +
                           os.path.join(TREE, "asan"))
    env["EMU_LOG"] = os.path.join(TREE, "emu_reports.log")
    env["EMU_DEFER"] = "1" if args.defer else "0"
    # test paths among the arguments replace the whole tests folder
    paths = [a for a in rest if a.startswith("tests")]
    cmd = [sys.executable, "-u", "-m", "pytest", *(paths or ["tests"]), "-m", "gpu", "-v",
           "-p", "no:cacheprovider", "-p", "no:faulthandler"]
    rest = [a for a in rest if a not in paths]
    if args.workers > 1:
        cmd += ["-n", str(args.workers)]
    rc = subprocess.call(cmd + rest, cwd=TREE, env=env)
    log = env["EMU_LOG"]
    if os.path.exists(log):
        print(open(log).read()[-4000:])
    sys.exit(rc)


if __name__ == "__main__":
    main()
