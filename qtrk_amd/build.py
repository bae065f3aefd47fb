"""Build recipe of libqtrk.so (gfx950) -- run by __graft_entry__.build().

The library is one shared object compiled by hipcc from qtrk_amd/csrc/ and
placed in-tree at qtrk_amd/libqtrk.so.  Every HIP source is compiled with
-ffp-contract=off (the kernels reproduce the reference CPU tracker's rounding,
which has no fused multiply-adds) and with packed-FP32 instruction selection
disabled.  The sources compile in parallel into qtrk_amd/_build/ and are then
linked; every call recompiles all of them.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libqtrk.so")
OBJDIR = os.path.join(HERE, "_build")

SOURCES = ["localize.hip", "beadfinder.hip", "lutmatch.hip", "lutmodel.hip", "hip_tracker.cpp", "c_api.cpp", "lutmatch_host.cpp",
           "lutmodel_host.cpp", "pipeline.cpp", "tracker_common.cpp", "result_manager.cpp", "noalloc.hip",
           "c_api_idle.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value",
    "-Wno-unused-result",
    "-pthread",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed: " + " ".join(cmd[-3:]))
    return r


def build_library(verbose=False, jobs=None):
    """Compiles every source into qtrk_amd/libqtrk.so (always rebuilds)."""
    os.makedirs(OBJDIR, exist_ok=True)
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
    objs, cmds = [], []
    for s in SOURCES:
        o = os.path.join(OBJDIR, os.path.splitext(s)[0] + ".o")
        objs.append(o)
        cmds.append([hipcc()] + HIPCC_FLAGS + inc + ["-c", os.path.join(CSRC, s), "-o", o])
    jobs = jobs or min(len(cmds), max(1, min(os.cpu_count() or 1, 16)))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread"] + objs + ["-o", tmp], verbose)
    os.replace(tmp, LIB)
    return LIB


CHECK_SRC = os.path.join(ROOT, "tests", "cxx_surface_check.cpp")
CHECK_BIN = os.path.join(ROOT, "tests", "_bin", "cxx_surface_check")
SPEED_SRC = os.path.join(HERE, "speed_test.c")
SPEED_BIN = os.path.join(HERE, "_bin", "qtrk_speed_test")


def build_checks(verbose=False):
    """The C++ check program of the tracker's C++ surface (tests/cxx_surface_check.cpp), linked
    against the built library; run by the GPU tests."""
    os.makedirs(os.path.dirname(CHECK_BIN), exist_ok=True)
    cxx = os.environ.get("CXX", "g++")
    _run([cxx, "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), CHECK_SRC, "-L" + HERE, "-lqtrk",
          "-Wl,-rpath,$ORIGIN/../../qtrk_amd", "-o", CHECK_BIN], verbose)
    return CHECK_BIN


def build_speed_test(verbose=False):
    """qtrk_speed_test (speed_test.c): the reference's speed-test clock through the C ABI, a plain C
    client of the built library; run by bench.py."""
    os.makedirs(os.path.dirname(SPEED_BIN), exist_ok=True)
    cc = os.environ.get("CC", "gcc")
    _run([cc, "-std=c11", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), SPEED_SRC, "-L" + HERE, "-lqtrk",
          "-Wl,-rpath,$ORIGIN/..", "-o", SPEED_BIN], verbose)
    return SPEED_BIN


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_checks(verbose=True))
    print(build_speed_test(verbose=True))
