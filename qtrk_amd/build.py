"""Build recipe of libqtrk.so (gfx950) -- run by __graft_entry__.build().

The library is one shared object compiled by hipcc from qtrk_amd/csrc/ and
placed in-tree at qtrk_amd/libqtrk.so.  Every HIP source is compiled with
-ffp-contract=off (the kernels reproduce the reference CPU tracker's rounding,
which has no fused multiply-adds) and with packed-FP32 instruction selection
disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libqtrk.so")

SOURCES = ["localize.hip", "beadfinder.hip", "lutmatch.hip", "lutmodel.hip", "hip_tracker.cpp", "c_api.cpp", "lutmatch_host.cpp",
           "lutmodel_host.cpp", "pipeline_kernels.hip", "pipeline.cpp", "tracker_common.cpp", "result_manager.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-shared",
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


def build_library(verbose=False):
    """Compiles every source into qtrk_amd/libqtrk.so (always rebuilds)."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + [
        "-I" + os.path.join(ROOT, "include"),
        "-I" + CSRC,
    ] + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libqtrk.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
