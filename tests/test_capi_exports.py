"""The C-ABI library loads on a machine without a GPU and exports every
function that include/lym.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "latticeyangmills_amd", "liblym.so")


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lym_\w+)\s*\(", text)))


@pytest.mark.parametrize("header", sorted(h for h in os.listdir(os.path.join(ROOT, "include")) if h.endswith(".h")))
def test_exports_every_declared_symbol(header):
    lib = ctypes.CDLL(LIB)
    names = declared(header)
    assert names, header
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_package_binding_loads():
    from latticeyangmills_amd import lib
    assert lib() is not None


def test_bindings_reject_wrong_shapes_without_a_gpu():
    """Shape checks raise before any call reaches the library (they are not asserts)."""
    import numpy as np
    from latticeyangmills_amd.capi import LatticeContext
    ctx = LatticeContext.__new__(LatticeContext)  # no device: only the Python-side checks run
    ctx.volume = 16
    with pytest.raises(ValueError, match="shape"):
        LatticeContext.upload(ctx, np.zeros((4, 15, 18)))
    with pytest.raises(ValueError, match="uniforms"):
        LatticeContext.metropolis(ctx, 6.0, 0.24, 2, 1, 0, "exph", 0, np.zeros(10))


def test_create_checks_the_lattice_before_the_device():
    """lym_create rejects bad extents, an unknown precision and lattices of
    44.7 M sites or more (48 V bytes per buffer descriptor must stay below
    2^31; 64^3 x 171 is the first of its shape beyond) with LYM_ERR_ARG before
    it looks for a GPU."""
    lib = ctypes.CDLL(LIB)
    lib.lym_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int]
    ctx = ctypes.c_void_p()
    for dims in [(64, 64, 64, 171), (1, 4, 4, 4), (4, 4, 4, 0)]:
        for prec in (0, 1):
            assert lib.lym_create(ctypes.byref(ctx), (ctypes.c_int * 4)(*dims), prec, 0) == -1, dims
    assert (1 << 31) // 48 < 64 * 64 * 64 * 171 and 64 * 64 * 64 * 170 < (1 << 31) // 48
    rc = lib.lym_create(ctypes.byref(ctx), (ctypes.c_int * 4)(2, 2, 2, 2), 7, 0)
    assert rc == -1  # unknown precision
