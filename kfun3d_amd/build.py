"""Builds the in-tree shared library kfun3d_amd/libkfun3d.so.

Host C (reader, preprocessing, solver driver) is compiled with gcc, the
device code with hipcc for gfx950; both without floating-point contraction
(-ffp-contract=off) so that the FP64 path rounds exactly as the reference.
Packed FP32 instructions are disabled for all device code.

    python -m kfun3d_amd.build          # library only
    python -m kfun3d_amd.build --all    # + oracle/ checker (+ oracle/_ref when the
                                        #   reference tree is present)
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libkfun3d.so")
ARCH = os.environ.get("KF_OFFLOAD_ARCH", "gfx950")

HOST_SRCS = ["gen.c", "msh.c", "setup.c", "plan.c", "aggregate.c", "solver.c"]
HIP_SRCS = ["ctx.hip", "flux.hip", "blas.hip", "fill.hip", "ilu.hip", "sptrsv.hip", "spmv.hip", "jacobi.hip", "coarse.hip",
            "mpnn.hip"]

CFLAGS = ["-O2", "-fPIC", "-std=gnu11", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wextra",
          "-Wno-unused-parameter", "-I" + os.path.join(ROOT, "include")]
HIPFLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
            "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed: " + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build_library(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    objs = []
    for s in HOST_SRCS:
        o = os.path.join(BUILD, s.replace(".c", ".o"))
        jobs.append(["gcc"] + CFLAGS + ["-c", os.path.join(HERE, "host", s), "-o", o])
        objs.append(o)
    for s in HIP_SRCS:
        o = os.path.join(BUILD, s.replace(".hip", ".o"))
        jobs.append(["hipcc"] + HIPFLAGS + ["-c", os.path.join(HERE, "csrc", s), "-o", o])
        objs.append(o)
    workers = min(8, int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run(["hipcc", "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lgomp", "-lm"])
    os.replace(tmp, LIB)
    return LIB


C_DRIVER_SRC = os.path.join(ROOT, "tests", "c", "kf_main.c")
C_DRIVER = os.path.join(ROOT, "tests", "c", "kf_main")


def build_c_driver():
    """tests/c/kf_main: a C program built only against include/kfun3d.h and
    linked to the library the way the reference's driver would be (run by
    the GPU tests; the rpath is relative, so the binary works wherever the
    tree is copied)."""
    tmp = C_DRIVER + ".tmp"
    _run(["gcc", "-O2", "-std=gnu11", "-Wall", "-Wextra", "-I" + os.path.join(ROOT, "include"), C_DRIVER_SRC,
          "-L" + HERE, "-lkfun3d", "-Wl,-rpath,$ORIGIN/../../kfun3d_amd", "-o", tmp])
    os.replace(tmp, C_DRIVER)
    return C_DRIVER


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"])
    if os.path.isdir("/root/reference/src"):
        _run(["make", "-s", "-j8", "-C", os.path.join(ROOT, "oracle"), "ref"])


if __name__ == "__main__":
    build_library(verbose="-v" in sys.argv)
    if "--all" in sys.argv:
        build_oracle()
    print(LIB)
