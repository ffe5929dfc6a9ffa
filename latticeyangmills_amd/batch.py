"""ctypes binding of the batched ensemble API (include/lym_batch.h)."""
import ctypes

import numpy as np

from .capi import CANDIDATES, LYM_OK, PRECISIONS, SCHED_LDS, SCHED_TWO_ROW, SCHEDULES, LymError, _D, _dptr, _links_ptrs, lib

SCHED_WHOLE_FLOW = 64  # LYM_SCHED_WHOLE_FLOW

_P = ctypes.c_void_p
_I = ctypes.c_int
_sigs_done = False


def _sigs(L):
    global _sigs_done
    if _sigs_done:
        return L
    c = ctypes
    L.lym_batch_create.argtypes = [c.POINTER(_P), _I, c.POINTER(_I), _I, _I]
    L.lym_batch_destroy.argtypes = [_P]
    L.lym_batch_last_error.argtypes = [_P]
    L.lym_batch_last_error.restype = c.c_char_p
    L.lym_batch_info.argtypes = [_P, c.POINTER(_I), c.POINTER(c.c_int64), c.POINTER(_I), c.POINTER(_I), c.POINTER(_I)]
    L.lym_batch_upload_links.argtypes = [_P, _I, c.POINTER(_D)]
    L.lym_batch_download_links.argtypes = [_P, _I, c.POINTER(_D)]
    L.lym_batch_set_hot.argtypes = [_P, c.POINTER(c.c_uint64)]
    L.lym_batch_metropolis.argtypes = [_P, c.c_double, c.c_double, _I, _I, _I, _I, c.POINTER(c.c_uint64),
                                       c.POINTER(c.c_int64)]
    L.lym_batch_observables.argtypes = [_P, _D]
    L.lym_batch_flow.argtypes = [_P, c.c_double, c.c_double, _I, _I, _D, _I, c.POINTER(_I)]
    L.lym_batch_set_schedule.argtypes = [_P, _I, _I, _I]
    L.lym_batch_set_barrier_timeout.argtypes = [_P, ctypes.c_double]
    L.lym_batch_autotune.argtypes = [_P, c.c_char_p, c.c_char_p, _I]
    _sigs_done = True
    return L


class BatchContext:
    """An ensemble of independent periodic lattices (a lym_batch)."""

    def __init__(self, dims_list, precision="mixed", device=0):
        self.dims = [tuple(int(x) for x in d) for d in dims_list]
        self.volumes = [int(np.prod(d)) for d in self.dims]
        self.nconf = len(self.dims)
        self._lib = _sigs(lib())
        h = _P()
        flat = (_I * (4 * self.nconf))(*[x for d in self.dims for x in d])
        rc = self._lib.lym_batch_create(ctypes.byref(h), self.nconf, flat, PRECISIONS[precision], device)
        if rc != LYM_OK:
            raise LymError(rc, "lym_batch_create", f"{self.nconf} lattices, precision={precision}")
        self._h = h

    def _check(self, rc, func):
        if rc != LYM_OK:
            raise LymError(rc, func, self._lib.lym_batch_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.lym_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, c, links):
        links = np.ascontiguousarray(links, dtype=np.float64)
        if links.shape != (4, self.volumes[c], 18):
            raise ValueError(f"links of lattice {c} must have shape (4, {self.volumes[c]}, 18), not {links.shape}")
        self._check(self._lib.lym_batch_upload_links(self._h, c, _links_ptrs(links)), "lym_batch_upload_links")

    def download(self, c):
        out = np.empty((4, self.volumes[c], 18))
        self._check(self._lib.lym_batch_download_links(self._h, c, _links_ptrs(out)), "lym_batch_download_links")
        return out

    def set_hot(self, seeds):
        s = (ctypes.c_uint64 * self.nconf)(*seeds)
        self._check(self._lib.lym_batch_set_hot(self._h, s), "lym_batch_set_hot")

    def metropolis(self, beta, eps, nhit, nsweeps, seeds, first_sweep=0, candidate="exph"):
        s = (ctypes.c_uint64 * self.nconf)(*seeds)
        acc = (ctypes.c_int64 * self.nconf)()
        self._check(self._lib.lym_batch_metropolis(self._h, beta, eps, nhit, nsweeps, first_sweep,
                                                   CANDIDATES[candidate], s, acc), "lym_batch_metropolis")
        return list(acc)

    def observables(self):
        out = np.empty((self.nconf, 4))
        self._check(self._lib.lym_batch_observables(self._h, _dptr(out)), "lym_batch_observables")
        return out

    def flow(self, eps, tau, measure_every=1, use_graph=True):
        rows = int(tau / eps)
        series = np.empty((self.nconf, max(rows, 1), 4))
        r = _I(0)
        self._check(self._lib.lym_batch_flow(self._h, eps, tau, measure_every, int(use_graph), _dptr(series),
                                             series.shape[1], ctypes.byref(r)), "lym_batch_flow")
        return series[:, : r.value]

    def set_schedule(self, grid=0, kind="column", target=4096, lds=False, two_row=False, whole_flow=False):
        code = (SCHEDULES[kind] | (SCHED_LDS if lds else 0) | (SCHED_TWO_ROW if two_row else 0)
                | (SCHED_WHOLE_FLOW if whole_flow else 0))
        self._check(self._lib.lym_batch_set_schedule(self._h, grid, code, target), "lym_batch_set_schedule")

    def set_barrier_timeout(self, seconds):
        self._check(self._lib.lym_batch_set_barrier_timeout(self._h, seconds), "lym_batch_set_barrier_timeout")

    def autotune(self, cache_path=None):
        buf = ctypes.create_string_buffer(256)
        self._check(self._lib.lym_batch_autotune(self._h, cache_path.encode() if cache_path else None, buf, 256),
                    "lym_batch_autotune")
        return buf.value.decode()

    def info(self):
        n, sites, grid, kind, target = _I(), ctypes.c_int64(), _I(), _I(), _I()
        self._check(self._lib.lym_batch_info(self._h, ctypes.byref(n), ctypes.byref(sites), ctypes.byref(grid),
                                             ctypes.byref(kind), ctypes.byref(target)), "lym_batch_info")
        return dict(nconf=n.value, total_sites=sites.value, grid=grid.value,
                    kind=[k for k, v in SCHEDULES.items() if v == kind.value & 3][0],
                    lds=bool(kind.value & SCHED_LDS), two_row=bool(kind.value & SCHED_TWO_ROW),
                    whole_flow=bool(kind.value & SCHED_WHOLE_FLOW), target=target.value)
