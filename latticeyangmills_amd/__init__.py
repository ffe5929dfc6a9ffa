"""MI355X-native hot path of LatticeYangMills: Wilson flow (RK3), clover
observables and checkerboard Metropolis as HIP kernels for gfx950 behind the
C ABI of include/lym.h.  The Python side is a ctypes binding for tests and
benchmarks; the C++ host classes live in latticeyangmills_amd/host."""
from .capi import LatticeContext, LymError, lib  # noqa: F401

__all__ = ["LatticeContext", "LymError", "lib"]
