"""Helpers for GPU tests: device buffers and the device-resident search entry
point, all through the C ABI (include/nndex_hip.h)."""

from __future__ import annotations

import ctypes

import numpy as np

from nndex_amd import _lib


def check(st):
    _lib.check(st)


class DevBuf:
    def __init__(self, nbytes: int):
        self.ptr = ctypes.c_void_p()
        self.nbytes = max(int(nbytes), 16)
        check(_lib.lib().nndex_device_malloc(ctypes.byref(self.ptr), self.nbytes))

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DevBuf":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        check(_lib.lib().nndex_memcpy_h2d(b.ptr, a.ctypes.data, a.nbytes))
        return b

    def to_array(self, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype)
        check(_lib.lib().nndex_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def offset(self, nbytes: int) -> ctypes.c_void_p:
        return ctypes.c_void_p(self.ptr.value + nbytes)

    def free(self):
        if self.ptr:
            check(_lib.lib().nndex_device_free(self.ptr))
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class BackendIndex:
    """nndex_gpu_index built from normalised padded storage (the gpu.rs boundary)."""

    def __init__(self, storage: np.ndarray, approx=False, enable_cache=False, device=-1):
        storage = np.ascontiguousarray(storage, np.float32)
        self.rows, self.pdims = storage.shape
        self.h = ctypes.c_void_p()
        check(_lib.lib().nndex_gpu_index_new(
            storage.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), self.rows, self.pdims,
            int(approx), int(enable_cache), device, ctypes.byref(self.h)))

    def search(self, q: np.ndarray, k: int, approx: bool = False):
        q = np.ascontiguousarray(q, np.float32)
        kk = min(k, self.rows)
        idx = np.zeros(kk, np.uint64)
        sim = np.zeros(kk, np.float32)
        n = ctypes.c_size_t()
        check(_lib.lib().nndex_gpu_index_search(
            self.h, q.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), kk, int(approx),
            idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
            sim.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(n)))
        return idx[: n.value], sim[: n.value]

    def search_batch(self, q: np.ndarray, k: int, approx: bool = False):
        q = np.ascontiguousarray(q, np.float32)
        nq = q.shape[0]
        kk = min(k, self.rows)
        idx = np.zeros((nq, kk), np.uint64)
        sim = np.zeros((nq, kk), np.float32)
        lens = np.zeros(nq, np.uintp)
        check(_lib.lib().nndex_gpu_index_search_batch(
            self.h, q.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), nq, kk, int(approx),
            idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
            sim.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
            lens.ctypes.data_as(ctypes.POINTER(ctypes.c_size_t))))
        return idx, sim, lens

    # ---- IVF ----
    def set_ivf(self, nlist=0, nprobe=0, iters=0):
        check(_lib.lib().nndex_gpu_index_set_ivf_params(self.h, nlist, nprobe, iters))

    def build_ivf(self) -> "_lib.IvfInfo":
        info = _lib.IvfInfo()
        check(_lib.lib().nndex_gpu_index_build_ivf(self.h, ctypes.byref(info)))
        return info

    def ivf_export(self):
        info = self.build_ivf()
        nl, pd = int(info.nlist), int(info.padded_dims)
        cent = np.zeros((max(nl, 1), pd), np.float32)
        off = np.zeros(nl + 1, np.uint64)
        orig = np.zeros(self.rows, np.uint64)
        assign = np.zeros(self.rows, np.uint64)
        u64 = ctypes.POINTER(ctypes.c_uint64)
        check(_lib.lib().nndex_gpu_index_ivf_export(
            self.h, cent.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
            off.ctypes.data_as(u64), orig.ctypes.data_as(u64), assign.ctypes.data_as(u64)))
        return info, cent[:nl], off, orig, assign

    def __del__(self):
        if getattr(self, "h", None):
            _lib.lib().nndex_gpu_index_free(self.h)
            self.h = None


class DeviceSearch:
    """Device-resident search with a private workspace and stream."""

    def __init__(self, handle, max_queries: int, max_k: int):
        self.handle = handle
        self.ws = ctypes.c_void_p()
        check(_lib.lib().nndex_workspace_new(handle, max_queries, max_k, ctypes.byref(self.ws)))
        self.stream = ctypes.c_void_p()
        check(_lib.lib().nndex_stream_create(ctypes.byref(self.stream)))

    def run_ivf(self, dq: DevBuf, nq: int, k: int, nprobe: int, didx: DevBuf, dsim: DevBuf,
                dlen: DevBuf, sync=True):
        check(_lib.lib().nndex_gpu_index_search_ivf_device(
            self.handle, self.ws, dq.ptr, nq, k, nprobe, didx.ptr, dsim.ptr, dlen.ptr,
            self.stream))
        if sync:
            check(_lib.lib().nndex_stream_synchronize(self.stream))

    def run(self, dq: DevBuf, nq: int, k: int, didx: DevBuf, dsim: DevBuf, sync=True):
        check(_lib.lib().nndex_gpu_index_search_device(self.handle, self.ws, dq.ptr, nq, k,
                                                       didx.ptr, dsim.ptr, self.stream))
        if sync:
            check(_lib.lib().nndex_stream_synchronize(self.stream))

    def __del__(self):
        L = _lib.lib()
        if getattr(self, "ws", None):
            L.nndex_workspace_free(self.ws)
        if getattr(self, "stream", None):
            L.nndex_stream_destroy(self.stream)


def gpu_available() -> bool:
    n = ctypes.c_int32(0)
    return _lib.lib().nndex_device_count(ctypes.byref(n)) == 0 and n.value > 0


def assert_topk_parity(idx, sim, fp_idx, fp_scores_full, tol=1e-5, tie=1e-6):
    """The parity bar: scores within `tol` of FP64; indices identical except
    where FP64 scores tie within `tie`."""
    idx = np.asarray(idx, np.int64)
    fp_idx = np.asarray(fp_idx, np.int64)
    assert len(idx) == len(fp_idx)
    np.testing.assert_allclose(np.asarray(sim, np.float64), fp_scores_full[idx], atol=tol, rtol=0)
    for pos, (a, b) in enumerate(zip(idx, fp_idx)):
        if a != b:
            assert abs(fp_scores_full[a] - fp_scores_full[b]) <= tie, (pos, a, b)
