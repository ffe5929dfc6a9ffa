"""Builds the native parts in-tree for gfx950.

liblym.so (HIP kernels + the C ABI of include/lym.h) and liblymhost.so (the
C++ host classes) land next to this file; the CPU oracle is built by
oracle/Makefile.  Run as ``python -m latticeyangmills_amd.build``.
"""
import concurrent.futures
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Every HIP translation unit is built without packed FP32 instructions, and
# with multiply-adds fused only where one source expression asks for it
# (-ffp-contract=on; the hot code spells its fusions out with fma()).  The
# default for HIP fuses across statements inside a basic block, so kernel
# variants that differ only in where a link is read from (LDS or HBM, other
# branches and barriers) fused different pairs and rounded differently.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=on",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

# One object per translation unit, compiled in parallel: the stage and
# Metropolis kernel families are instantiated per link precision in units of
# their own (lym_launch.hpp), the C ABI of the two paths in lym_capi / lym_batch.
LIB_SOURCES = ["lym_bflow_f64.hip", "lym_bflow_f32.hip", "lym_bflow_f32r2.hip", "lym_bstage_f64.hip",
               "lym_bstage_f32.hip", "lym_bstage_f32r2.hip", "lym_stage_f64.hip", "lym_stage_f32.hip",
               "lym_stage_f32r2.hip", "lym_metro_f64.hip", "lym_metro_f32.hip", "lym_capi.hip", "lym_batch.hip",
               "lym_noalloc.hip"]
# The batch kernels that loop over schedule slots (the persistent grid of a
# stage, the whole-flow launch) would have machine-level loop-invariant code
# motion hoist the constants of the stage body out of those loops and hold
# them in registers across them (145-156 VGPRs, 3 waves per SIMD, against
# 91-96 and 5 without it), so the batch units are built without it; the
# one-workgroup-per-slot kernels in the same units take the same time either way.
UNIT_FLAGS = {s: ["-mllvm", "-disable-machine-licm"] for s in LIB_SOURCES
              if s.startswith("lym_bflow_") or s.startswith("lym_bstage_")}
OBJ_DIR = os.path.join(PKG, "_build")


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def _jobs():
    n = os.cpu_count() or 1
    return max(1, min(n, int(os.environ.get("MAX_JOBS", n)), len(LIB_SOURCES)))


def build_lym(extra=()):
    out = os.path.join(PKG, "liblym.so")
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs = [os.path.join(OBJ_DIR, os.path.splitext(s)[0] + ".o") for s in LIB_SOURCES]

    def compile_one(i):
        _run([HIPCC, *HIP_FLAGS, *UNIT_FLAGS.get(LIB_SOURCES[i], []), *extra, "-c", "-I", INCLUDE, "-I", CSRC,
              "-o", objs[i], os.path.join(CSRC, LIB_SOURCES[i])])

    with concurrent.futures.ThreadPoolExecutor(_jobs()) as pool:
        for f in [pool.submit(compile_one, i) for i in range(len(LIB_SOURCES))]:
            f.result()
    _run([HIPCC, *HIP_FLAGS, *extra, "-shared", "-o", out + ".tmp", *objs])
    os.replace(out + ".tmp", out)
    return out


HOST = os.path.join(PKG, "host")
HOST_SOURCES = ["math.cpp", "devicefield.cpp", "gpuclasses.cpp", "confio.cpp", "jsonlite.cpp", "scale.cpp"]
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
