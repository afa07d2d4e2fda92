"""CPU oracle for nndex's exact search path -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper over ``oracle/liborc.so`` (built from ``nndex_oracle.c`` by
``oracle/Makefile``).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker or as the timed CPU baseline; the product package ``nndex_amd`` never
imports it.

Every wrapper mirrors one C function whose header cites the reference
file:line it restates (fblissjr/nndex v0.2.1).
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")

ORC_OK, ORC_NON_FINITE, ORC_ZERO_NORM = 0, 1, 2
SEARCH_STRATEGIES = ["Serial", "Gemv", "ParallelScores", "ParallelChunked", "ParallelFold"]
BATCH_STRATEGIES = ["Gemm", "GemmChunked", "MatrixChunked", "QueryParallel", "ProfiledPerQuery"]


class OracleError(RuntimeError):
    def __init__(self, code: int):
        self.code = code
        name = {ORC_NON_FINITE: "NonFiniteInput", ORC_ZERO_NORM: "ZeroNorm"}.get(code, str(code))
        super().__init__(name)
        self.kind = name


def build() -> str:
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None
_P = ctypes.c_void_p
_SZ = ctypes.c_size_t
_F = ctypes.POINTER(ctypes.c_float)
_D = ctypes.POINTER(ctypes.c_double)
_U64 = ctypes.POINTER(ctypes.c_uint64)
_I = ctypes.c_int


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_ensure_finite": (_I, [_F, _SZ]),
            "orc_pad_rows": (None, [_F, _SZ, _SZ, _SZ, _F]),
            "orc_normalize_rows": (_I, [_F, _SZ, _SZ, _SZ, _F]),
            "orc_normalize_vector": (_I, [_F, _SZ, _SZ, _F]),
            "orc_dot_product": (ctypes.c_float, [_F, _F, _SZ]),
            "orc_dot_product_simd": (ctypes.c_float, [_F, _F, _SZ]),
            "orc_dot_unrolled_8": (ctypes.c_float, [_F, _F, _SZ]),
            "orc_dot_ndarray": (ctypes.c_float, [_F, _F, _SZ]),
            "orc_topk_new": (_P, [_SZ]),
            "orc_topk_free": (None, [_P]),
            "orc_topk_push": (None, [_P, ctypes.c_uint64, ctypes.c_float]),
            "orc_topk_push_slow": (None, [_P, ctypes.c_uint64, ctypes.c_float]),
            "orc_topk_merge": (None, [_P, _P]),
            "orc_topk_min_threshold": (ctypes.c_float, [_P]),
            "orc_topk_len": (_SZ, [_P]),
            "orc_topk_into_sorted": (_SZ, [_P, _U64, _F]),
            "orc_topk_from_scores": (_SZ, [_F, _SZ, _SZ, _SZ, _U64, _F]),
            "orc_profile_cutoff": (_SZ, [_SZ, _SZ, _SZ]),
            "orc_cpu_search_strategy": (_I, [_SZ, _SZ, _SZ, _SZ, _SZ, _SZ, _I]),
            "orc_cpu_batch_strategy": (_I, [_SZ, _SZ, _SZ, _I]),
            "orc_parallel_cutoff": (_SZ, [_SZ, _I]),
            "orc_chunk_rows": (_SZ, [_SZ, _SZ]),
            "orc_batch_chunk_rows": (_SZ, [_SZ]),
            "orc_cpu_search": (_SZ, [_F, _SZ, _SZ, _F, _SZ, _I, _U64, _F, ctypes.POINTER(_I)]),
            "orc_gpu_exact_search": (_SZ, [_F, _SZ, _SZ, _F, _SZ, _U64, _F]),
            "orc_cpu_search_batch": (
                None,
                [_F, _SZ, _SZ, _F, _SZ, _SZ, _I, _U64, _F, ctypes.POINTER(_SZ), ctypes.POINTER(_I)],
            ),
            "orc_sgemm_scores": (None, [_F, _SZ, _F, _SZ, _SZ, _I, _F]),
            "orc_fp64_topk": (_I, [_F, _SZ, _SZ, _F, _SZ, _SZ, _I, _I, _U64, _D]),
            "orc_fp64_scores": (_I, [_F, _SZ, _SZ, _F, _I, _I, _D]),
            "orc_isqrt": (_SZ, [_SZ]),
            "orc_choose_nlist": (_SZ, [_SZ]),
            "orc_choose_nprobe": (_SZ, [_SZ, _SZ, _SZ]),
            "orc_pick_medoid_row": (_SZ, [_F, _SZ, _SZ]),
            "orc_kmeans_pp_rows": (None, [_F, _SZ, _SZ, _SZ, _I, _U64]),
            "orc_ivf_build": (_P, [_F, _SZ, _SZ, _SZ, _SZ, _I]),
            "orc_ivf_free": (None, [_P]),
            "orc_ivf_nlist": (_SZ, [_P]),
            "orc_ivf_provides_speedup": (_I, [_P]),
            "orc_ivf_centroids": (_F, [_P]),
            "orc_ivf_offsets": (_U64, [_P]),
            "orc_ivf_original_indices": (_U64, [_P]),
            "orc_ivf_assignments": (_U64, [_P]),
            "orc_ivf_search": (_SZ, [_P, _F, _SZ, _SZ, _U64, _F]),
            "orc_ivf_probe": (_SZ, [_P, _F, _SZ, _U64]),
            "orc_ivf_search_gpu": (_SZ, [_P, _F, _F, _SZ, _SZ, _U64, _F]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _fp(a: np.ndarray):
    return a.ctypes.data_as(_F)


def _check(code: int) -> None:
    if code != ORC_OK:
        raise OracleError(code)


# ---- normalisation (lib.rs:433-599) ----

def padded_dims(dims: int) -> int:
    return (dims + 7) // 8 * 8


def ensure_finite(values) -> None:
    v = _f32(values).ravel()
    _check(lib().orc_ensure_finite(_fp(v), v.size))


def pad_rows(values, rows: int, dims: int, pdims: int) -> np.ndarray:
    v = _f32(values).ravel()
    out = np.empty(rows * pdims, np.float32)
    lib().orc_pad_rows(_fp(v), rows, dims, pdims, _fp(out))
    return out.reshape(rows, pdims)


def normalize_rows(values, rows: int, dims: int, pdims: int | None = None) -> np.ndarray:
    pdims = padded_dims(dims) if pdims is None else pdims
    v = _f32(values).ravel()
    out = np.zeros(rows * pdims, np.float32)
    _check(lib().orc_normalize_rows(_fp(v), rows, dims, pdims, _fp(out)))
    return out.reshape(rows, pdims)


def normalize_vector(values, pdims: int | None = None) -> np.ndarray:
    v = _f32(values).ravel()
    pdims = padded_dims(v.size) if pdims is None else pdims
    out = np.zeros(pdims, np.float32)
    _check(lib().orc_normalize_vector(_fp(v), v.size, pdims, _fp(out)))
    return out


def dot_product(a, b) -> float:
    a, b = _f32(a), _f32(b)
    return float(lib().orc_dot_product(_fp(a), _fp(b), min(a.size, b.size)))


def dot_unrolled_8(a, b) -> float:
    a, b = _f32(a), _f32(b)
    return float(lib().orc_dot_unrolled_8(_fp(a), _fp(b), min(a.size, b.size)))


def dot_ndarray(a, b) -> float:
    a, b = _f32(a), _f32(b)
    return float(lib().orc_dot_ndarray(_fp(a), _fp(b), min(a.size, b.size)))


# ---- top-k (topk.rs) ----

class TopK:
    """TopKAccumulator (topk.rs:44-152)."""

    def __init__(self, k: int):
        self.k = k
        self._h = lib().orc_topk_new(k)

    def push(self, index: int, sim: float) -> None:
        lib().orc_topk_push(self._h, index, sim)

    def push_slow(self, index: int, sim: float) -> None:
        lib().orc_topk_push_slow(self._h, index, sim)

    def merge(self, other: "TopK") -> None:
        lib().orc_topk_merge(self._h, other._h)
        other._h = None

    @property
    def min_threshold(self) -> float:
        return float(lib().orc_topk_min_threshold(self._h))

    def into_sorted(self):
        n = max(self.k, 1)
        idx = np.zeros(n, np.uint64)
        sim = np.zeros(n, np.float32)
        got = lib().orc_topk_into_sorted(self._h, idx.ctypes.data_as(_U64), _fp(sim))
        self._h = None
        return idx[:got], sim[:got]

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_topk_free(self._h)


def topk_from_scores(scores, k: int, row_offset: int = 0):
    s = _f32(scores).ravel()
    n = max(min(k, s.size), 1)
    idx = np.zeros(n, np.uint64)
    sim = np.zeros(n, np.float32)
    got = lib().orc_topk_from_scores(_fp(s), s.size, k, row_offset, idx.ctypes.data_as(_U64),
                                     _fp(sim))
    return idx[:got], sim[:got]


# ---- strategy selection (profiles.rs) ----

def profile_cutoff(rows: int, dims: int, query_rows: int) -> int:
    return int(lib().orc_profile_cutoff(rows, dims, query_rows))


def cpu_search_strategy(profile_cut, rows, dims, par_cut, chunk, scores_max, has_blas) -> str:
    return SEARCH_STRATEGIES[lib().orc_cpu_search_strategy(profile_cut, rows, dims, par_cut,
                                                           chunk, scores_max, int(has_blas))]


def cpu_batch_strategy(rows, dims, query_rows, prefer_query_parallel) -> str:
    return BATCH_STRATEGIES[lib().orc_cpu_batch_strategy(rows, dims, query_rows,
                                                         int(prefer_query_parallel))]


def chunk_rows(rows: int, dims: int) -> int:
    return int(lib().orc_chunk_rows(rows, dims))


def batch_chunk_rows(dims: int) -> int:
    return int(lib().orc_batch_chunk_rows(dims))


def parallel_cutoff(dims: int, threads: int) -> int:
    return int(lib().orc_parallel_cutoff(dims, threads))


# ---- exact CPU index (cpu.rs) over normalised padded storage ----

def cpu_search(matrix: np.ndarray, query: np.ndarray, k: int, threads: int = 0):
    m = _f32(matrix)
    q = _f32(query).ravel()
    rows, dims = m.shape
    kk = max(min(k, rows), 1)
    idx = np.zeros(kk, np.uint64)
    sim = np.zeros(kk, np.float32)
    st = _I(0)
    got = lib().orc_cpu_search(_fp(m), rows, dims, _fp(q), k, threads, idx.ctypes.data_as(_U64),
                               _fp(sim), ctypes.byref(st))
    return idx[:got], sim[:got], SEARCH_STRATEGIES[st.value]


def gpu_exact_search(matrix: np.ndarray, query: np.ndarray, k: int):
    """The GPU backend's exact search (src/gpu.rs:575-589): every row, in row
    order, pushed into one bounded heap (its own choice among tied rows)."""
    m = _f32(matrix)
    q = _f32(query).ravel()
    rows, dims = m.shape
    kk = max(min(k, rows), 1)
    idx = np.zeros(kk, np.uint64)
    sim = np.zeros(kk, np.float32)
    got = lib().orc_gpu_exact_search(_fp(m), rows, dims, _fp(q), k, idx.ctypes.data_as(_U64),
                                     _fp(sim))
    return idx[:got], sim[:got]


def cpu_search_batch(matrix: np.ndarray, queries: np.ndarray, k: int, threads: int = 0):
    m = _f32(matrix)
    q = _f32(queries)
    rows, dims = m.shape
    nq = q.shape[0]
    kk = max(min(k, rows), 1)
    idx = np.zeros((nq, kk), np.uint64)
    sim = np.zeros((nq, kk), np.float32)
    lens = np.zeros(nq, np.uintp)
    st = _I(0)
    lib().orc_cpu_search_batch(_fp(m), rows, dims, _fp(q), nq, k, threads,
                               idx.ctypes.data_as(_U64), _fp(sim),
                               lens.ctypes.data_as(ctypes.POINTER(_SZ)), ctypes.byref(st))
    return idx, sim, lens, BATCH_STRATEGIES[st.value]


def sgemm_scores(queries: np.ndarray, matrix: np.ndarray, threads: int = 0) -> np.ndarray:
    q, m = _f32(queries), _f32(matrix)
    out = np.empty((q.shape[0], m.shape[0]), np.float32)
    lib().orc_sgemm_scores(_fp(q), q.shape[0], _fp(m), m.shape[0], m.shape[1], threads, _fp(out))
    return out


# ---- FP64 oracle ----

def fp64_topk(matrix, queries, k: int, normalized: bool, threads: int = 0):
    m = _f32(matrix)
    q = _f32(queries)
    if q.ndim == 1:
        q = q[None, :]
    rows, dims = m.shape
    nq = q.shape[0]
    kk = min(k, rows)
    idx = np.zeros((nq, kk), np.uint64)
    sim = np.zeros((nq, kk), np.float64)
    _check(lib().orc_fp64_topk(_fp(m), rows, dims, _fp(q), nq, k, int(normalized), threads,
                               idx.ctypes.data_as(_U64), sim.ctypes.data_as(_D)))
    return idx, sim


def fp64_scores(matrix, query, normalized: bool, threads: int = 0) -> np.ndarray:
    m = _f32(matrix)
    q = _f32(query).ravel()
    out = np.empty(m.shape[0], np.float64)
    _check(lib().orc_fp64_scores(_fp(m), m.shape[0], m.shape[1], _fp(q), int(normalized), threads,
                                 out.ctypes.data_as(_D)))
    return out


# ---- IVF (approx.rs) ----

def isqrt(n: int) -> int:
    return int(lib().orc_isqrt(n))


def choose_nlist(rows: int) -> int:
    return int(lib().orc_choose_nlist(rows))


def choose_nprobe(nlist: int, k: int, rows: int) -> int:
    return int(lib().orc_choose_nprobe(nlist, k, rows))


def pick_medoid_row(matrix) -> int:
    """pick_medoid_row (src/approx.rs:409-429)."""
    m = _f32(matrix)
    return int(lib().orc_pick_medoid_row(_fp(m), m.shape[0], m.shape[1]))


def kmeans_pp_rows(matrix, nlist: int, threads: int = 0) -> np.ndarray:
    """Rows the k-means++ seeding copies (src/approx.rs:357-406)."""
    m = _f32(matrix)
    out = np.zeros(nlist, np.uint64)
    lib().orc_kmeans_pp_rows(_fp(m), m.shape[0], m.shape[1], nlist, threads,
                             out.ctypes.data_as(_U64))
    return out


class Ivf:
    """ApproxIndex (approx.rs:11-300) over normalised padded storage."""

    def __init__(self, matrix: np.ndarray, nlist: int = 0, iters: int = 0, threads: int = 0):
        self._m = _f32(matrix)
        self.rows, self.dims = self._m.shape
        self._h = lib().orc_ivf_build(_fp(self._m), self.rows, self.dims, nlist, iters, threads)
        self.nlist = int(lib().orc_ivf_nlist(self._h))

    @property
    def provides_speedup(self) -> bool:
        return bool(lib().orc_ivf_provides_speedup(self._h))

    def centroids(self) -> np.ndarray:
        p = lib().orc_ivf_centroids(self._h)
        if not p:
            return np.zeros((0, self.dims), np.float32)
        return np.ctypeslib.as_array(p, shape=(self.nlist * self.dims,)).reshape(
            self.nlist, self.dims).copy()

    def offsets(self) -> np.ndarray:
        return np.ctypeslib.as_array(lib().orc_ivf_offsets(self._h), shape=(self.nlist + 1,)).copy()

    def original_indices(self) -> np.ndarray:
        return np.ctypeslib.as_array(lib().orc_ivf_original_indices(self._h),
                                     shape=(self.rows,)).copy()

    def assignments(self) -> np.ndarray:
        return np.ctypeslib.as_array(lib().orc_ivf_assignments(self._h), shape=(self.rows,)).copy()

    def search(self, query, k: int, nprobe: int = 0):
        q = _f32(query).ravel()
        kk = max(min(k, self.rows), 1)
        idx = np.zeros(kk, np.uint64)
        sim = np.zeros(kk, np.float32)
        got = lib().orc_ivf_search(self._h, _fp(q), k, nprobe, idx.ctypes.data_as(_U64), _fp(sim))
        return idx[:got], sim[:got]

    def search_gpu(self, query, k: int, nprobe: int = 0):
        """The GPU backend's approximate search (src/gpu.rs:448-462)."""
        q = _f32(query).ravel()
        kk = max(min(k, self.rows), 1)
        idx = np.zeros(kk, np.uint64)
        sim = np.zeros(kk, np.float32)
        got = lib().orc_ivf_search_gpu(self._h, _fp(self._m), _fp(q), k, nprobe,
                                       idx.ctypes.data_as(_U64), _fp(sim))
        return idx[:got], sim[:got]

    def probe(self, query, nprobe: int) -> np.ndarray:
        q = _f32(query).ravel()
        out = np.zeros(self.nlist, np.uint64)
        got = lib().orc_ivf_probe(self._h, _fp(q), nprobe, out.ctypes.data_as(_U64))
        return out[:got]

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_ivf_free(self._h)
            self._h = None
