"""Build recipe of libgcr.so (HIP kernels + C-ABI) for gfx950.

    python -m gcr_amd.build        # from gliding-collective-rotation_amd/

The library is built in-tree (gcr_amd/libgcr.so) so it travels with the
source tree to the machine that runs it.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)  # gliding-collective-rotation_amd/
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
INCLUDE = os.path.join(REPO, "include")
LIB = os.path.join(PKG, "libgcr.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 VALU ops are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
    "-pthread",
]


def build(verbose=False):
    srcs = [os.path.join(CSRC, "gcr_capi.hip"), os.path.join(CSRC, "gcr_init.cpp"),
            os.path.join(CSRC, "gcr_snapshot.cpp")]
    cmd = [HIPCC, *FLAGS, "-I", CSRC, "-I", INCLUDE, *srcs, "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
