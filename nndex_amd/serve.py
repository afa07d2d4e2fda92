"""Serving loop over a resident GPU index (C ABI ``nndex_server_*``).

Continuous batching: requests are queued and executors (one HIP stream,
workspace and staging each) batch every pending request of the same kind
(exact, or IVF at one nprobe) as they free up; the number of executors is
admitted by HBM footprint.  No reference counterpart: the reference serves one
call at a time (src/gpu.rs:511-592).
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

EXACT = 0
IVF = 1

_FP = ctypes.POINTER(ctypes.c_float)
_U64P = ctypes.POINTER(ctypes.c_uint64)


class Server:
    """Serve searches on a backend handle (``nndex_gpu_index*``).

    Queries are normalised, zero-padded rows (``padded_dims`` columns), as at
    the backend boundary.  ``graphs=True`` replays captured HIP graphs for exact
    batches above 4 rows (rows rounded up to a multiple of 16 up to 128 and of
    64 above, k to a power of two).  ``submit`` returns a ticket at once; ``wait``
    returns ``(indices[uint64, rows x k], similarities[float32, rows x k],
    counts[uint64, rows], latency_ms)``.
    """

    def __init__(self, backend_handle, padded_dims: int, rows: int, slots: int = 4,
                 max_batch_queries: int = 2048, max_k: int = 100, ivf_nprobe_max: int = 0,
                 hbm_budget_bytes: int = 0, graphs: bool = False):
        self._L = _lib.lib()
        self.pdims = int(padded_dims)
        self.rows = int(rows)
        opt = _lib.ServerOptions(slots, max_batch_queries, max_k, ivf_nprobe_max,
                                 hbm_budget_bytes, 1 if graphs else 0)
        self._h = ctypes.c_void_p()
        _lib.check(self._L.nndex_server_new(backend_handle, ctypes.byref(opt),
                                            ctypes.byref(self._h)))
        self._meta = {}

    def submit(self, queries: np.ndarray, k: int, mode: int = EXACT, nprobe: int = 0) -> int:
        q = np.ascontiguousarray(queries, np.float32)
        if q.ndim == 1:
            q = q[None, :]
        if q.shape[1] != self.pdims:
            raise ValueError(f"queries need {self.pdims} columns, got {q.shape[1]}")
        t = ctypes.c_uint64()
        _lib.check(self._L.nndex_server_submit(self._h, q.ctypes.data_as(_FP), q.shape[0], k,
                                               mode, nprobe, ctypes.byref(t)))
        self._meta[t.value] = (q.shape[0], min(k, self.rows))
        return t.value

    def submit_raw(self, queries: np.ndarray, k: int, mode: int = EXACT, nprobe: int = 0,
                   normalize: bool = True) -> int:
        """Raw rows (``dims`` columns, padding to the index's width) checked
        and, with ``normalize``, normalised on the device as the front-end's
        batch search does; a failed check fails this request at ``wait``."""
        q = np.ascontiguousarray(queries, np.float32)
        if q.ndim == 1:
            q = q[None, :]
        t = ctypes.c_uint64()
        _lib.check(self._L.nndex_server_submit_raw(self._h, q.ctypes.data_as(_FP), q.shape[0],
                                                   q.shape[1], int(bool(normalize)), k, mode,
                                                   nprobe, ctypes.byref(t)))
        self._meta[t.value] = (q.shape[0], min(k, self.rows))
        return t.value

    def wait(self, ticket: int):
        nq, k = self._meta.pop(ticket)
        idx = np.zeros((nq, k), np.uint64)
        sim = np.zeros((nq, k), np.float32)
        lens = np.zeros(nq, np.uint64)
        lat = ctypes.c_double()
        _lib.check(self._L.nndex_server_wait(self._h, ticket, idx.ctypes.data_as(_U64P),
                                             sim.ctypes.data_as(_FP), lens.ctypes.data_as(_U64P),
                                             ctypes.byref(lat)))
        return idx, sim, lens, lat.value

    def stats(self) -> dict:
        st = _lib.ServerStats()
        _lib.check(self._L.nndex_server_get_stats(self._h, ctypes.byref(st)))
        return st.to_dict()

    def reset_stats(self) -> None:
        _lib.check(self._L.nndex_server_reset_stats(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.nndex_server_free(self._h)
            self._h = None

    def __del__(self):
        self.close()
