"""ctypes binding of liblym.so (the C ABI declared in include/lym.h).

There is no CPU fallback: if the HIP library is missing, importing the
binding raises, and every call reports the library's own error.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblym.so")

LYM_OK = 0
ERRORS = {-1: "invalid argument", -2: "HIP runtime error", -3: "device out of memory", -4: "no GPU visible"}
PRECISIONS = {"fp64": 0, "mixed": 1}
SCHEDULES = {"linear": 0, "contig": 1, "column": 2}
SCHED_LDS = 16  # or-ed into a schedule kind: stage each run's own links in LDS
SCHED_TWO_ROW = 32  # or-ed into a schedule kind (mixed): the flow keeps links in two-row form
CANDIDATES = {"su2": 0, "exph": 1}
UNIFORMS_PER_HIT = {"su2": 13, "exph": 9}

_lib = None
_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.c_int


class LymError(RuntimeError):
    def __init__(self, code, func, msg):
        super().__init__(f"{func} failed with {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


def _sig(lib):
    c = ctypes
    lib.lym_create.argtypes = [c.POINTER(_P), c.POINTER(_I), _I, _I]
    lib.lym_destroy.argtypes = [_P]
    lib.lym_last_error.argtypes = [_P]
    lib.lym_last_error.restype = c.c_char_p
    lib.lym_info.argtypes = [_P, c.POINTER(_I), c.POINTER(c.c_int64), c.POINTER(_I)]
    lib.lym_upload_links.argtypes = [_P, c.POINTER(_D)]
    lib.lym_download_links.argtypes = [_P, c.POINTER(_D)]
    lib.lym_set_cold.argtypes = [_P]
    lib.lym_set_hot.argtypes = [_P, c.c_uint64]
    lib.lym_staples.argtypes = [_P, _I, _D]
    lib.lym_derivative.argtypes = [_P, _I, _D]
    lib.lym_observables.argtypes = [_P, _D]
    lib.lym_flow_steps.argtypes = [_P, c.c_double, _I]
    lib.lym_flow.argtypes = [_P, c.c_double, c.c_double, _I, _D, _I, c.POINTER(_I)]
    lib.lym_set_graph.argtypes = [_P, _I]
    lib.lym_metropolis.argtypes = [_P, c.c_double, c.c_double, _I, _I, _I, _I, c.c_uint64, _D,
                                   c.POINTER(c.c_int64)]
    lib.lym_synchronize.argtypes = [_P]
    lib.lym_check_links.argtypes = [_P, c.POINTER(c.c_double), c.POINTER(c.c_int64)]
    lib.lym_profile_flow.argtypes = [_P, c.c_double, _I, _I, _D, _D]
    lib.lym_copy_bandwidth.argtypes = [_I, c.c_int64, _I, _D]
    lib.lym_set_schedule.argtypes = [_P, _I, _I]
    lib.lym_autotune.argtypes = [_P, c.c_char_p, c.c_char_p, _I]
    return lib


def copy_bandwidth(nbytes=1 << 30, reps=20, device=0):
    """GB/s of a streaming device copy (read + write bytes over kernel time)."""
    gbs = ctypes.c_double(0)
    rc = lib().lym_copy_bandwidth(device, nbytes, reps, ctypes.byref(gbs))
    if rc != LYM_OK:
        raise LymError(rc, "lym_copy_bandwidth", "")
    return gbs.value


def lib():
    """The loaded liblym.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing; build it with `python -m latticeyangmills_amd.build`")
        _lib = _sig(ctypes.CDLL(LIB_PATH))
    return _lib


def _dptr(a):
    return a.ctypes.data_as(_D)


def _links_ptrs(arr):
    return (_D * 4)(*[_dptr(arr[mu]) for mu in range(4)])


class LatticeContext:
    """One periodic lattice resident on a GPU (a lym_ctx).

    Link arrays on the host are numpy float64 of shape (4, V, 18): the
    reference's GluonField, four Lattice<SU3> of V sites with 18 reals each.
    """

    def __init__(self, dims, precision="fp64", device=0):
        self.dims = tuple(int(d) for d in dims)
        self.volume = int(np.prod(self.dims))
        self.precision = precision
        self._lib = lib()
        h = _P()
        d = (_I * 4)(*self.dims)
        rc = self._lib.lym_create(ctypes.byref(h), d, PRECISIONS[precision], device)
        if rc != LYM_OK:
            raise LymError(rc, "lym_create", f"dims={self.dims} precision={precision}")
        self._h = h

    def _check(self, rc, func):
        if rc != LYM_OK:
            raise LymError(rc, func, self._lib.lym_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.lym_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- field transfer ---------------------------------------------------
    def upload(self, links):
        links = np.ascontiguousarray(links, dtype=np.float64)
        if links.shape != (4, self.volume, 18):
            raise ValueError(f"links must have shape (4, {self.volume}, 18), not {links.shape}")
        self._check(self._lib.lym_upload_links(self._h, _links_ptrs(links)), "lym_upload_links")

    def download(self):
        out = np.empty((4, self.volume, 18))
        self._check(self._lib.lym_download_links(self._h, _links_ptrs(out)), "lym_download_links")
        return out

    def set_cold(self):
        self._check(self._lib.lym_set_cold(self._h), "lym_set_cold")

    def set_hot(self, seed):
        self._check(self._lib.lym_set_hot(self._h, seed), "lym_set_hot")

    # --- action -----------------------------------------------------------
    def staples(self, mu):
        out = np.empty((self.volume, 18))
        self._check(self._lib.lym_staples(self._h, mu, _dptr(out)), "lym_staples")
        return out

    def derivative(self, mu):
        out = np.empty((self.volume, 18))
        self._check(self._lib.lym_derivative(self._h, mu, _dptr(out)), "lym_derivative")
        return out

    # --- observables and flow ----------------------------------------------
    def observables(self):
        """(plaq, energy_raw, topc, topc_scale) as SuperObs::compute."""
        out = np.empty(4)
        self._check(self._lib.lym_observables(self._h, _dptr(out)), "lym_observables")
        return tuple(float(x) for x in out)

    def flow_steps(self, eps, nsteps):
        self._check(self._lib.lym_flow_steps(self._h, eps, nsteps), "lym_flow_steps")

    def flow(self, eps, tau, measure_every=1):
        rows = int(tau / eps)
        series = np.empty((max(rows, 1), 4))
        r = _I(0)
        self._check(self._lib.lym_flow(self._h, eps, tau, measure_every, _dptr(series), series.shape[0],
                                       ctypes.byref(r)), "lym_flow")
        return series[: r.value]

    def set_graph(self, on):
        """flow() records its launches once as a HIP graph and replays it."""
        self._check(self._lib.lym_set_graph(self._h, 1 if on else 0), "lym_set_graph")

    def synchronize(self):
        self._check(self._lib.lym_synchronize(self._h), "lym_synchronize")

    def check_links(self):
        """(max unitarity / determinant deviation, links with non-finite entries)"""
        dev, bad = ctypes.c_double(0), ctypes.c_int64(0)
        self._check(self._lib.lym_check_links(self._h, ctypes.byref(dev), ctypes.byref(bad)), "lym_check_links")
        return dev.value, bad.value

    def profile_flow(self, eps, nsteps, measure_every=0):
        total = ctypes.c_double(0)
        st = np.zeros(4)
        self._check(self._lib.lym_profile_flow(self._h, eps, nsteps, measure_every, ctypes.byref(total),
                                               _dptr(st)), "lym_profile_flow")
        return total.value, st

    # --- work schedule ----------------------------------------------------
    def set_schedule(self, kind, target=0, lds=False, two_row=False):
        code = SCHEDULES[kind] | (SCHED_LDS if lds else 0) | (SCHED_TWO_ROW if two_row else 0)
        self._check(self._lib.lym_set_schedule(self._h, code, target), "lym_set_schedule")

    def autotune(self, cache_path=None):
        buf = ctypes.create_string_buffer(256)
        self._check(self._lib.lym_autotune(self._h, cache_path.encode() if cache_path else None, buf, 256),
                    "lym_autotune")
        return buf.value.decode()

    # --- Metropolis -------------------------------------------------------
    def metropolis(self, beta, eps, nhit, nsweeps, first_sweep=0, candidate="exph", seed=0, predrawn=None):
        acc = ctypes.c_int64(0)
        pd = None
        if predrawn is not None:
            pd = np.ascontiguousarray(predrawn, dtype=np.float64)
            need = nsweeps * 4 * self.volume * nhit * UNIFORMS_PER_HIT[candidate]
            if pd.size != need:
                raise ValueError(f"predrawn must hold nsweeps * 4V * nhit * {UNIFORMS_PER_HIT[candidate]} = {need} "
                                 f"uniforms, not {pd.size}")
        self._check(self._lib.lym_metropolis(self._h, beta, eps, nhit, nsweeps, first_sweep, CANDIDATES[candidate],
                                             seed, _dptr(pd) if pd is not None else None, ctypes.byref(acc)),
                    "lym_metropolis")
        return acc.value
