"""Builds the native parts in-tree for gfx950.

liblym.so (HIP kernels + the C ABI of include/lym.h) and liblymhost.so (the
C++ host classes) land next to this file; the CPU oracle is built by
oracle/Makefile.  Run as ``python -m latticeyangmills_amd.build``.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Every HIP translation unit is built without packed FP32 instructions.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

LIB_SOURCES = ["lym_capi.hip", "lym_batch.hip"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_lym(extra=()):
    out = os.path.join(PKG, "liblym.so")
    srcs = [os.path.join(CSRC, s) for s in LIB_SOURCES]
    _run([HIPCC, *HIP_FLAGS, *extra, "-shared", "-I", INCLUDE, "-I", CSRC, "-o", out + ".tmp", *srcs])
    os.replace(out + ".tmp", out)
    return out


def build_oracle(with_ref=None):
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"])
    ref = os.environ.get("LYM_REFERENCE", "/root/reference")
    if with_ref is None:
        with_ref = os.path.isfile(os.path.join(ref, "src", "Math", "su3.cpp"))
    if with_ref:
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", "REF=" + ref])


def build_all():
    build_lym()
    build_oracle()


if __name__ == "__main__":
    extra = sys.argv[1:]
    build_lym(extra)
    build_oracle()
