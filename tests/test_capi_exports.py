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


def test_null_handles_are_refused_without_a_gpu():
    """Every entry point that takes a context or a batch refuses a null one
    with LYM_ERR_ARG (destroy accepts it as a no-op) before it touches a
    device, and no call throws across the ABI; checked on a machine with no
    GPU, so a call that reached the runtime first would report another code."""
    c = ctypes
    lib = c.CDLL(LIB)
    P, I, D, L = c.c_void_p, c.c_int, c.c_double, c.c_int64
    buf = (c.c_double * 64)()
    ptrs = (c.c_void_p * 4)(*[c.cast(buf, c.c_void_p)] * 4)
    i4 = (c.c_int * 4)(4, 4, 4, 4)
    n = c.c_int()
    calls = {
        "lym_destroy": ([P], [None], 0),
        "lym_info": ([P, P, P, P], [None, i4, None, None], -1),
        "lym_upload_links": ([P, P], [None, ptrs], -1),
        "lym_download_links": ([P, P], [None, ptrs], -1),
        "lym_set_cold": ([P], [None], -1),
        "lym_set_hot": ([P, c.c_uint64], [None, 1], -1),
        "lym_staples": ([P, I, P], [None, 0, buf], -1),
        "lym_derivative": ([P, I, P], [None, 0, buf], -1),
        "lym_observables": ([P, P], [None, buf], -1),
        "lym_flow_steps": ([P, D, I], [None, 0.01, 1], -1),
        "lym_flow": ([P, D, D, I, P, I, P], [None, 0.01, 0.1, 1, buf, 16, c.byref(n)], -1),
        "lym_set_graph": ([P, I], [None, 1], -1),
        "lym_metropolis": ([P, D, D, I, I, I, I, c.c_uint64, P, P], [None, 6.0, 0.24, 1, 1, 0, 1, 1, None, None], -1),
        "lym_check_links": ([P, P, P], [None, buf, None], -1),
        "lym_synchronize": ([P], [None], -1),
        "lym_profile_flow": ([P, D, I, I, P, P], [None, 0.01, 1, 1, buf, buf], -1),
        "lym_set_schedule": ([P, I, I], [None, 2, 4096], -1),
        "lym_autotune": ([P, P, P, I], [None, None, None, 0], -1),
        "lym_batch_destroy": ([P], [None], 0),
        "lym_batch_info": ([P, P, P, P, P, P], [None, None, None, None, None, None], -1),
        "lym_batch_upload_links": ([P, I, P], [None, 0, ptrs], -1),
        "lym_batch_download_links": ([P, I, P], [None, 0, ptrs], -1),
        "lym_batch_set_hot": ([P, P], [None, buf], -1),
        "lym_batch_metropolis": ([P, D, D, I, I, I, I, P, P], [None, 6.0, 0.24, 1, 1, 0, 1, buf, None], -1),
        "lym_batch_observables": ([P, P], [None, buf], -1),
        "lym_batch_flow": ([P, D, D, I, I, P, I, P], [None, 0.01, 0.1, 1, 1, buf, 16, c.byref(n)], -1),
        "lym_batch_set_schedule": ([P, I, I, I], [None, 0, 2, 4096], -1),
        "lym_batch_set_barrier_timeout": ([P, D], [None, 1.0], -1),
        "lym_batch_autotune": ([P, P, P, I], [None, None, None, 0], -1),
    }
    for name, (argtypes, args, want) in calls.items():
        f = getattr(lib, name)
        f.argtypes, f.restype = argtypes, I
        assert f(*args) == want, name
