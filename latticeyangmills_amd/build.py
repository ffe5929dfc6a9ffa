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


HOST = os.path.join(PKG, "host")
HOST_SOURCES = ["math.cpp", "gpuclasses.cpp", "confio.cpp", "jsonlite.cpp", "scale.cpp"]
CXX = os.environ.get("CXX", "g++")
CXX_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-Wall", "-Wno-unused-result"]


def build_host():
    """liblymhost.so (the C++ classes over the C ABI) and the lym_app CLI."""
    inc = ["-I", INCLUDE, "-I", os.path.join(HOST, "include")]
    link = ["-L", PKG, "-llym", "-Wl,-rpath,$ORIGIN"]
    out = os.path.join(PKG, "liblymhost.so")
    srcs = [os.path.join(HOST, "src", s) for s in HOST_SOURCES]
    _run([CXX, *CXX_FLAGS, *inc, "-shared", "-o", out + ".tmp", *srcs, *link])
    os.replace(out + ".tmp", out)
    app = os.path.join(PKG, "lym_app")
    _run([CXX, *CXX_FLAGS, *inc, "-o", app + ".tmp", os.path.join(HOST, "src", "lym_app.cpp"),
          "-L", PKG, "-llymhost", *link])
    os.replace(app + ".tmp", app)


def build_oracle(with_ref=None):
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"])
    ref = os.environ.get("LYM_REFERENCE", "/root/reference")
    if with_ref is None:
        with_ref = os.path.isfile(os.path.join(ref, "src", "Math", "su3.cpp"))
    if with_ref:
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", "REF=" + ref])


def build_all():
    build_lym()
    build_host()
    build_oracle()


if __name__ == "__main__":
    extra = sys.argv[1:]
    build_lym(extra)
    build_host()
    build_oracle()
