"""lym_observables_noalloc (include/lym.h): the observables of a link field the
caller holds in device memory, computed with the library's static scratch
array instead of any allocation (the test allocates the caller's buffers
through the HIP runtime the library itself runs on).  Checked through the C ABI against
lym_observables on the same links (plaquette and <E> to 1e-12 relative, Q to
1e-12 of its cancellation scale), on two calls in a row of which the second
has the smaller lattice, in both precisions; sizes beyond the maximum are
rejected before any device work."""
import ctypes

import numpy as np
import pytest

from latticeyangmills_amd import LatticeContext, lib
from latticeyangmills_amd.capi import PRECISIONS
from tests import oracle_lib as orc

pytestmark = pytest.mark.gpu

LYM_ERR_TOO_LARGE = -5
MAX_SITES = 1 << 24


def _fn():
    L = lib()
    f = L.lym_observables_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    f.restype = ctypes.c_int
    return f


def _hip():
    """The HIP runtime liblym.so runs on (its path from the process's mappings),
    for the caller-side buffers of the test."""
    lib()
    paths = sorted({line.split()[-1] for line in open("/proc/self/maps")
                    if "libamdhip64.so" in line and "torch" not in line})
    assert paths, "libamdhip64 not mapped"
    h = ctypes.CDLL(paths[0])
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


H2D, D2H = 1, 2


def device_copy(h, host):
    """host numpy array -> a new device buffer (the caller's allocation)"""
    p = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(p), host.nbytes) == 0
    assert h.hipMemcpy(p, host.ctypes.data, host.nbytes, H2D) == 0
    return p


def noalloc_observables(links, dims, precision):
    """(4, V, 18) host links -> the library's device layout [mu][e][s] of
    (re, im) pairs in a device buffer, then lym_observables_noalloc on the
    null stream into a device buffer of 4 doubles."""
    h = _hip()
    V = links.shape[1]
    soa = np.ascontiguousarray(links.reshape(4, V, 9, 2).transpose(0, 2, 1, 3))
    if precision == "mixed":
        soa = soa.astype(np.float32)
    field = device_copy(h, soa)
    out = device_copy(h, np.full(4, np.nan))
    try:
        rc = _fn()(field, (ctypes.c_int * 4)(*dims), PRECISIONS[precision], out, None)
        assert rc == 0, lib().lym_last_error(None)
        assert h.hipDeviceSynchronize() == 0
        res = np.empty(4)
        assert h.hipMemcpy(res.ctypes.data, out, res.nbytes, D2H) == 0
    finally:
        h.hipFree(field)
        h.hipFree(out)
    return res


def ctx_observables(links, dims, precision):
    """lym_observables of the uploaded links, and the links as the context
    stores them (FP32 fields derive row 2 from the rounded rows 0, 1)"""
    with LatticeContext(dims, precision) as ctx:
        ctx.upload(links)
        return np.array(ctx.observables()), ctx.download()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_matches_lym_observables_twice(precision):
    big, small = (8, 8, 8, 16), (4, 4, 6, 8)
    for dims, seed in [(big, 71), (small, 72)]:
        want, stored = ctx_observables(orc.hot_start(dims, seed), dims, precision)
        got = noalloc_observables(stored, dims, precision)
        assert abs(got[0] / want[0] - 1) < 1e-12, (dims, got, want)
        assert abs(got[1] / want[1] - 1) < 1e-12, (dims, got, want)
        assert abs(got[2] - want[2]) < 1e-12 * want[3], (dims, got, want)
        assert abs(got[3] / want[3] - 1) < 1e-12, (dims, got, want)


def test_noalloc_rejects_beyond_maximum():
    f = _fn()
    lib().lym_last_error.restype = ctypes.c_char_p
    dims = (256, 256, 256, 2)  # 2^25 sites
    dummy = ctypes.c_void_p(16)  # never dereferenced: the size check comes first
    rc = f(dummy, (ctypes.c_int * 4)(*dims), PRECISIONS["fp64"], dummy, None)
    assert rc == LYM_ERR_TOO_LARGE
    assert b"maximum" in lib().lym_last_error(None)
    assert int(np.prod(dims)) > MAX_SITES
