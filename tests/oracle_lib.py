"""ctypes binding of the CPU oracle (oracle/lym_oracle.c) for the tests.

Test infrastructure: the oracle is the checker, never the thing measured.
Link arrays are numpy float64 (4, V, 18), the reference's GluonField layout.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liblym_oracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref_su3.so")

_D = ctypes.POINTER(ctypes.c_double)
_I4 = ctypes.c_int * 4
_lib = None
_ref = None


def _p(a):
    return a.ctypes.data_as(_D)


def lib():
    global _lib
    if _lib is None:
        if not os.path.isfile(ORACLE_SO):
            raise RuntimeError("oracle not built: make -C oracle")
        L = ctypes.CDLL(ORACLE_SO)
        L.or_su3_retr_mul.restype = ctypes.c_double
        L.or_su3_exp.restype = ctypes.c_int
        L.or_flow_step.restype = ctypes.c_int
        L.or_flow.restype = ctypes.c_int
        L.or_flow.argtypes = [_D, _I4, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, _D]
        L.or_flow_step.argtypes = [_D, _I4, ctypes.c_double]
        L.or_flow_step_lattice.restype = ctypes.c_int
        L.or_flow_step_lattice.argtypes = [_D, _I4, ctypes.c_double]
        L.or_observables.argtypes = [_D, _I4, ctypes.c_int, _D]
        L.or_staples.argtypes = [_D, _I4, ctypes.c_int, _D]
        L.or_derivative.argtypes = [_D, _I4, ctypes.c_int, _D]
        L.or_hot_start.argtypes = [_D, _I4, ctypes.c_uint64]
        L.or_metropolis.restype = ctypes.c_int
        L.or_metropolis.argtypes = [_D, _I4, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _D, ctypes.POINTER(ctypes.c_int64)]
        L.or_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32)] * 3
        L.or_rand_su3_transf.argtypes = [_D, ctypes.c_double, _D]
        L.or_rand_su3_exph.argtypes = [_D, ctypes.c_double, _D]
        _lib = L
    return _lib


def ref():
    """The reference's own su3.cpp (built into oracle/_ref), or None."""
    global _ref
    if _ref is None and os.path.isfile(REF_SO):
        R = ctypes.CDLL(REF_SO)
        R.ref_su3_retr_mul.restype = ctypes.c_double
        _ref = R
    return _ref


def volume(dims):
    return int(np.prod(dims))


# ---- SU(3) helpers -----------------------------------------------------------
def su3_mul(a, b):
    r = np.empty(18)
    lib().or_su3_mul(_p(np.ascontiguousarray(a, float)), _p(np.ascontiguousarray(b, float)), _p(r))
    return r


def su3_exp(q):
    r = np.empty(18)
    flag = lib().or_su3_exp(_p(np.ascontiguousarray(q, float)), _p(r))
    return r, flag


def to18(M):
    out = np.empty(M.shape[:-2] + (18,))
    out[..., 0::2] = M.real.reshape(M.shape[:-2] + (9,))
    out[..., 1::2] = M.imag.reshape(M.shape[:-2] + (9,))
    return out


def from18(a):
    a = np.asarray(a)
    return (a[..., 0::2] + 1j * a[..., 1::2]).reshape(a.shape[:-1] + (3, 3))


# ---- lattice-wide -------------------------------------------------------------
def hot_start(dims, seed):
    links = np.empty((4, volume(dims), 18))
    lib().or_hot_start(_p(links), _I4(*dims), seed)
    return links


def cold_start(dims):
    links = np.zeros((4, volume(dims), 18))
    links[..., 0] = links[..., 8] = links[..., 16] = 1.0
    return links


def observables(links, dims, sum_mode=0):
    out = np.empty(4)
    lib().or_observables(_p(np.ascontiguousarray(links)), _I4(*dims), sum_mode, _p(out))
    return out


def staples(links, dims, mu):
    out = np.empty((volume(dims), 18))
    lib().or_staples(_p(np.ascontiguousarray(links)), _I4(*dims), mu, _p(out))
    return out


def derivative(links, dims, mu):
    out = np.empty((volume(dims), 18))
    lib().or_derivative(_p(np.ascontiguousarray(links)), _I4(*dims), mu, _p(out))
    return out


def flow_step(links, dims, eps):
    links = np.array(links, copy=True)
    assert lib().or_flow_step(_p(links), _I4(*dims), eps) == 0
    return links


def flow_step_lattice(links, dims, eps):
    """or_flow_step in the reference's whole-lattice-temporary structure"""
    links = np.array(links, copy=True)
    assert lib().or_flow_step_lattice(_p(links), _I4(*dims), eps) == 0
    return links


def flow(links, dims, eps, tau, measure_every=1, sum_mode=0):
    links = np.array(links, copy=True)
    rows = int(tau / eps)
    series = np.empty((rows, 4))
    r = lib().or_flow(_p(links), _I4(*dims), eps, tau, measure_every, sum_mode, _p(series))
    assert r == rows
    return series, links


def metropolis(links, dims, beta, eps, nhit, nsweeps, first_sweep=0, candidate=1, seed=0, predrawn=None):
    links = np.array(links, copy=True)
    acc = ctypes.c_int64(0)
    pd = None if predrawn is None else np.ascontiguousarray(predrawn, float)
    rc = lib().or_metropolis(_p(links), _I4(*dims), beta, eps, nhit, nsweeps, first_sweep, candidate, seed,
                             _p(pd) if pd is not None else None, ctypes.byref(acc))
    assert rc == 0
    return links, acc.value


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return list(o)


def thermalised(dims, beta=6.0, sweeps=10, seed=1, eps=0.24, nhit=10):
    """Hot start plus Metropolis sweeps, all by the oracle."""
    links = hot_start(dims, seed)
    links, _ = metropolis(links, dims, beta, eps, nhit, sweeps, 0, 1, seed)
    return links
