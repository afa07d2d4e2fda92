"""nndex_amd -- MI355X (gfx950) search backend for nndex.

Python surface of the reference's PyO3 module ``nndex._nndex``
(src/python_bindings.rs:30-297): ``NNdex(data, normalized=False, approx=False,
backend="cpu", enable_cache=True, gpu_device_index=None)`` with ``from_file``,
``search(query, k=10, dataframe=None)``, the ``backend``/``rows``/``dims``/
``gpu_info`` properties, and ``list_gpu_devices()``.

All compute runs in libnndex_hip.so (hand-written HIP kernels behind the C ABI
of include/nndex_hip.h).  This build carries the GPU backend only, like the
reference built without its ``cpu`` feature: ``backend="cpu"`` raises
``ValueError("cpu backend is unavailable ...")`` (src/lib.rs:244-254).
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import NNdexStatusError

__all__ = ["NNdex", "list_gpu_devices", "NNdexStatusError"]
__version__ = "0.1.0"

_BACKENDS = {"cpu": 0, "gpu": 1}


def _parse_backend(raw: str) -> int:
    """parse_backend, src/python_bindings.rs:300-308 (case-insensitive)."""
    low = str(raw).lower()
    if low not in _BACKENDS:
        raise ValueError(f"unsupported backend '{low}', expected one of: cpu, gpu")
    return _BACKENDS[low]


def _dataframe_kind(value) -> str | None:
    """detect_dataframe_kind, src/python_bindings.rs:403-417."""
    t = type(value)
    if t.__name__ == "DataFrame":
        mod = t.__module__ or ""
        if "pandas" in mod:
            return "pandas"
        if "polars" in mod:
            return "polars"
    return None


def _parse_input(value) -> tuple[np.ndarray, int, int]:
    """parse_input, src/python_bindings.rs:311-400: numpy f32/f64, lists,
    ``.to_numpy()`` objects -> (flat f32 values, rows, dims)."""
    if hasattr(value, "to_numpy") and not isinstance(value, np.ndarray):
        conv = value.to_numpy()
        if conv is not value:
            return _parse_input(conv)
    if isinstance(value, np.ndarray):
        if value.dtype not in (np.float32, np.float64):
            value = value.astype(np.float32)
        arr = np.ascontiguousarray(value, dtype=np.float32)
        if arr.ndim == 1:
            return arr, 1, arr.shape[0]
        if arr.ndim == 2:
            return arr.reshape(-1), arr.shape[0], arr.shape[1]
        raise ValueError("only 1D and 2D arrays are supported for inputs")
    if isinstance(value, (list, tuple)):
        if len(value) and isinstance(value[0], (list, tuple, np.ndarray)):
            if len(value[0]) == 0:
                raise ValueError("input matrix must be non-empty")
            dims = len(value[0])
            if any(len(r) != dims for r in value):
                raise ValueError("2D list input must have consistent row lengths")
            arr = np.asarray(value, dtype=np.float32)
            return arr.reshape(-1), arr.shape[0], dims
        if len(value) == 0:
            raise ValueError("input vector must be non-empty")
        arr = np.asarray(value, dtype=np.float32)
        return arr, 1, arr.shape[0]
    raise ValueError("unsupported input type: expected list, numpy array, "
                     "pandas/polars Series, or DataFrame")


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _u64ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


class NNdex:
    """In-memory cosine-similarity index (src/python_bindings.rs:30-237)."""

    def __init__(self, data, normalized: bool = False, approx: bool = False,
                 backend: str = "cpu", enable_cache: bool = True,
                 gpu_device_index: int | None = None):
        opts = _lib.IndexOptions(
            normalized=int(bool(normalized)), approx=int(bool(approx)),
            backend=_parse_backend(backend), enable_cache=int(bool(enable_cache)),
            gpu_device_index=-1 if gpu_device_index is None else int(gpu_device_index))
        values, rows, dims = _parse_input(data)
        self._build(values, rows, dims, opts)

    def _build(self, values: np.ndarray, rows: int, dims: int, opts) -> None:
        L = _lib.lib()
        handle = ctypes.c_void_p()
        # Construction errors map to ValueError (src/python_bindings.rs:247).
        st = L.nndex_new(_fptr(values), values.size, rows, dims, ctypes.byref(opts),
                         ctypes.byref(handle))
        if st != _lib.OK:
            raise ValueError(L.nndex_last_error().decode())
        self._h = handle
        self._keep = values  # the constructor cache keys on the source pointer
        self._rows = rows
        self._dims = dims

    @classmethod
    def from_file(cls, path: str, *, key: str | None = None, normalized: bool = False,
                  approx: bool = False, backend: str = "cpu", enable_cache: bool = True,
                  gpu_device_index: int | None = None) -> "NNdex":
        """from_file (src/python_bindings.rs:68-86) over the loaders of
        nndex_amd.io (src/python_io.rs)."""
        from .io import load_matrix_from_disk

        opts = _lib.IndexOptions(
            normalized=int(bool(normalized)), approx=int(bool(approx)),
            backend=_parse_backend(backend), enable_cache=int(bool(enable_cache)),
            gpu_device_index=-1 if gpu_device_index is None else int(gpu_device_index))
        try:
            matrix = load_matrix_from_disk(path, key)
        except (OSError, ValueError, KeyError) as e:
            raise ValueError(str(e)) from None
        self = cls.__new__(cls)
        self._build(np.ascontiguousarray(matrix, np.float32).reshape(-1), matrix.shape[0],
                    matrix.shape[1], opts)
        return self

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.lib().nndex_free(h)
            self._h = None

    @property
    def backend(self) -> str:
        return "gpu" if _lib.lib().nndex_backend_of(self._h) == 1 else "cpu"

    @property
    def rows(self) -> int:
        return self._rows

    @property
    def dims(self) -> int:
        return self._dims

    @property
    def backend_handle(self):
        """The ``nndex_gpu_index*`` behind a GPU index (for the device entry
        points and the serving loop of ``nndex_amd.serve``)."""
        h = _lib.lib().nndex_backend_handle(self._h)
        return ctypes.c_void_p(h) if h else None

    @property
    def padded_dims(self) -> int:
        return (self._dims + 7) // 8 * 8

    def autotune(self, reps: int = 20, batch: bool = True) -> str:
        """Time the launch shapes of this GPU index's kernels on its own data
        and keep the fastest: the single-query scan (64-row tiles per
        wavefront, load flags) and, with ``batch``, the batched MFMA scan's
        tile shapes per scoring precision.  Returns the timing report.
        Results do not depend on the shapes chosen."""
        h = self.backend_handle
        if h is None:
            raise RuntimeError("autotune needs a GPU index")
        buf = ctypes.create_string_buffer(1 << 16)
        _lib.check(_lib.lib().nndex_gpu_index_autotune(h, reps, buf, 1 << 16))
        report = buf.value.decode()
        if batch:
            _lib.check(_lib.lib().nndex_gpu_index_autotune_batch(h, 3, buf, 1 << 16))
            report += buf.value.decode()
        return report

    # ---- IVF persistence (the reference rebuilds ApproxIndex on every
    # construction, src/gpu.rs:435-441; these keep a GPU build) -----------

    def _gpu_handle(self, what: str):
        h = self.backend_handle
        if h is None:
            raise RuntimeError(f"{what} needs a GPU index")
        return h

    def ivf_export(self) -> dict:
        """The IVF build of this index (built now if needed) as numpy arrays:
        ``centroids`` [nlist, padded_dims] f32, ``offsets`` [nlist + 1],
        ``original_indices`` [rows] (reordered position -> row) and
        ``assignments`` [rows], all uint64 (ApproxIndex, src/approx.rs:11-25)."""
        h = self._gpu_handle("ivf_export")
        L = _lib.lib()
        info = _lib.IvfInfo()
        _lib.check(L.nndex_gpu_index_build_ivf(h, ctypes.byref(info)))
        nl, pd = int(info.nlist), int(info.padded_dims)
        cent = np.zeros((max(nl, 1), pd), np.float32)
        off = np.zeros(nl + 1, np.uint64)
        orig = np.zeros(self._rows, np.uint64)
        assign = np.zeros(self._rows, np.uint64)
        _lib.check(L.nndex_gpu_index_ivf_export(h, _fptr(cent), _u64ptr(off), _u64ptr(orig),
                                                _u64ptr(assign)))
        return {"centroids": cent[:nl], "offsets": off, "original_indices": orig,
                "assignments": assign}

    def ivf_import(self, centroids, offsets, original_indices, assignments=None) -> None:
        """Install a saved IVF build (the arrays of ``ivf_export``) instead of
        running the k-means build; the index must not have built one yet."""
        h = self._gpu_handle("ivf_import")
        cent = np.ascontiguousarray(centroids, np.float32)
        if cent.ndim != 2 or cent.shape[1] != self.padded_dims:
            raise ValueError(f"centroids must be [nlist, {self.padded_dims}] (padded dims)")
        nl = cent.shape[0]
        off = np.ascontiguousarray(offsets, np.uint64)
        orig = np.ascontiguousarray(original_indices, np.uint64)
        if off.shape != (nl + 1,) or orig.shape != (self._rows,):
            raise ValueError("offsets must hold nlist + 1 values and original_indices rows")
        asg = None
        if assignments is not None:
            asg = np.ascontiguousarray(assignments, np.uint64)
            if asg.shape != (self._rows,):
                raise ValueError("assignments must hold one value per row")
        _lib.check(_lib.lib().nndex_gpu_index_ivf_import(
            h, nl, _fptr(cent), _u64ptr(off), _u64ptr(orig),
            _u64ptr(asg) if asg is not None else None, None))

    def save_ivf(self, path: str) -> None:
        """Write ``ivf_export()`` to an ``.npz`` file."""
        np.savez(path, rows=np.uint64(self._rows), padded_dims=np.uint64(self.padded_dims),
                 **self.ivf_export())

    def load_ivf(self, path: str) -> None:
        """Install an IVF build written by ``save_ivf`` for the same rows."""
        with np.load(path) as f:
            if int(f["rows"]) != self._rows or int(f["padded_dims"]) != self.padded_dims:
                raise ValueError("saved IVF index was built for a different shape "
                                 f"({int(f['rows'])} x {int(f['padded_dims'])})")
            self.ivf_import(f["centroids"], f["offsets"], f["original_indices"],
                            f["assignments"])

    @property
    def gpu_info(self) -> dict | None:
        info = _lib.GpuDeviceInfo()
        has = ctypes.c_int32(0)
        _lib.check(_lib.lib().nndex_gpu_info(self._h, ctypes.byref(info), ctypes.byref(has)))
        return info.to_dict() if has.value else None

    # ---- search ---------------------------------------------------------

    def _search_single(self, q: np.ndarray, k: int):
        kk = max(min(k, self._rows), 1)
        idx = np.zeros(kk, np.uint64)
        sim = np.zeros(kk, np.float32)
        n = ctypes.c_size_t(0)
        st = _lib.lib().nndex_search(self._h, _fptr(q), q.size, k, _u64ptr(idx), _fptr(sim),
                                     ctypes.byref(n))
        if st != _lib.OK:
            raise RuntimeError(_lib.lib().nndex_last_error().decode())
        return idx[: n.value], sim[: n.value]

    def _search_batch(self, q: np.ndarray, rows: int, k: int):
        kk = max(min(k, self._rows), 1)
        idx = np.zeros((rows, kk), np.uint64)
        sim = np.zeros((rows, kk), np.float32)
        lens = np.zeros(rows, np.uintp)
        st = _lib.lib().nndex_search_batch(
            self._h, _fptr(q), q.size, rows, k, _u64ptr(idx), _fptr(sim),
            lens.ctypes.data_as(ctypes.POINTER(ctypes.c_size_t)))
        if st != _lib.OK:
            raise RuntimeError(_lib.lib().nndex_last_error().decode())
        # Rectangular output padded with (0, 0.0) past each row's length
        # (batch_neighbors_to_numpy, src/python_bindings.rs:464-497).
        cols = int(lens[0]) if rows else 0
        for r in range(rows):
            if int(lens[r]) < cols:
                idx[r, int(lens[r]):] = 0
                sim[r, int(lens[r]):] = 0.0
        return idx[:, :cols], sim[:, :cols]

    def search(self, query, k: int = 10, dataframe=None):
        """search (src/python_bindings.rs:125-236): 1-D query -> (uint64[k],
        float32[k]); 2-D -> (uint64[Q,k], float32[Q,k]); a 1-row 2-D query is
        answered by the single-query path."""
        if k < 0:
            raise OverflowError("can't convert negative int to unsigned")
        out_df = self._resolve_output_dataframe(dataframe)
        if isinstance(query, np.ndarray) and query.dtype == np.float32 and query.ndim in (1, 2):
            values = np.ascontiguousarray(query)
            rows = 1 if values.ndim == 1 else values.shape[0]
            dims = values.shape[-1]
            values = values.reshape(-1)
        else:
            values, rows, dims = _parse_input(query)
        if dims != self._dims:
            raise ValueError(f"query dims mismatch: expected {self._dims}, got {dims}")
        if rows == 1:
            res = self._search_single(values, k)
            if out_df is not None:
                return self._to_dataframe([res], out_df)
            return res
        res = self._search_batch(values, rows, k)
        if out_df is not None:
            return self._to_dataframe(list(zip(res[0], res[1])), out_df)
        return res

    def _resolve_output_dataframe(self, dataframe):
        """resolve_output_dataframe, src/python_bindings.rs:420-438."""
        if dataframe is None:
            return None
        kind = _dataframe_kind(dataframe)
        if kind is None:
            raise ValueError("`dataframe` must be a pandas.DataFrame or polars.DataFrame")
        n = dataframe.shape[0] if kind == "pandas" else dataframe.height
        if n != self._rows:
            raise ValueError("provided dataframe row count must match index rows "
                             f"(index rows={self._rows}, dataframe rows={n})")
        return kind, dataframe

    @staticmethod
    def _to_dataframe(batch, out_df):
        """neighbors_to_source_dataframe, src/python_bindings.rs:499-553."""
        kind, src = out_df
        frames = []
        for idx, sim in batch:
            rows = [int(i) for i in idx]
            sims = [float(s) for s in sim]
            if kind == "pandas":
                frames.append(src.take(rows).assign(similarity=sims))
            else:
                import polars as pl

                frames.append(src[rows].with_columns(pl.Series("similarity", sims)))
        return frames[0] if len(frames) == 1 else frames


def list_gpu_devices() -> list[dict]:
    """list_gpu_devices (src/python_bindings.rs:273-289)."""
    L = _lib.lib()
    count = ctypes.c_size_t(0)
    _lib.check(L.nndex_list_gpu_devices(None, 0, ctypes.byref(count)))
    n = count.value
    if n == 0:
        return []
    arr = (_lib.GpuDeviceInfo * n)()
    _lib.check(L.nndex_list_gpu_devices(arr, n, ctypes.byref(count)))
    return [arr[i].to_dict() for i in range(min(n, count.value))]
