"""Builds the tsgpu shared library in-tree with hipcc for gfx950.

``python -m tskit_amd.build`` compiles every source under ``tskit_amd/csrc``
into ``tskit_amd/libtsgpu.so``.  The library exposes a C ABI (see
``include/tsgpu.h``) and is loaded with ctypes by ``tskit_amd._lib``.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtsgpu.so")
ARCH = os.environ.get("TSG_OFFLOAD_ARCH", "gfx950")

SOURCES = [
    "tsg_kernels.hip",
    "tsg_relate.hip",
    "tsg_noalloc.hip",
    "tsg_relvec.hip",
    "tsg_api.cpp",
    "tsg_treeseq.cpp",
    "tsg_sim.cpp",
]

FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    f"--offload-arch={ARCH}",
    # summary values must round identically wherever they are evaluated
    "-ffp-contract=off",
    # no packed FP32 instructions (v_pk_*_f32)
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the tsgpu library needs ROCm's hipcc")
    return path


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *FLAGS, "-I", CSRC, "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(
        line
        for line in res.stderr.splitlines()
        if "packed-fp32-ops" not in line and "not a recognized feature" not in line
    )
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
