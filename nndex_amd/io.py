"""Disk loaders for NNdex.from_file (src/python_io.rs:80-377).

``.npy`` (2-D f32), ``.npz`` (2-D f32 under ``key``) and ``.parquet`` (a
list<f32> or fixed_size_list<f32> column named ``key``) become a row-major
float32 matrix.  Error texts follow DiskLoadError (src/python_io.rs:20-74).
Arrays are read with ``allow_pickle=False``.

A ``.npy`` file, and an ``.npz`` member stored without compression (what
``numpy.savez`` writes), are memory-mapped rather than read: construction then
copies the rows from the file's pages into pinned staging buffers and on to
the device in chunks, and no host copy of the whole matrix is made (the
reference reads the file into a Vec first, src/python_io.rs:112-165).
"""

from __future__ import annotations

import os
import zipfile

import numpy as np


class DiskLoadError(ValueError):
    pass


def _as_matrix(path: str, arr: np.ndarray) -> np.ndarray:
    """matrix_from_ndarray, src/python_io.rs:238-262."""
    if arr.ndim != 2:
        raise DiskLoadError(f"failed to read matrix from '{path}': expected a 2-D array, "
                            f"found {arr.ndim}-D")
    if arr.dtype != np.float32:
        raise DiskLoadError(f"failed to read matrix from '{path}': expected float32, "
                            f"found {arr.dtype}")
    if arr.shape[0] == 0 or arr.shape[1] == 0:
        raise DiskLoadError(f"matrix loaded from '{path}' is empty or has zero dimensions")
    return np.ascontiguousarray(arr)


def load_matrix_from_disk(path: str, key: str | None = None) -> np.ndarray:
    """load_matrix_from_disk, src/python_io.rs:80-110."""
    ext = os.path.splitext(path)[1].lower().lstrip(".")
    if ext == "npy":
        try:
            arr = np.load(path, mmap_mode="r", allow_pickle=False)
        except (OSError, ValueError) as e:
            raise DiskLoadError(f"failed to read .npy matrix from '{path}': {e}") from None
        return _as_matrix(path, arr)
    if ext == "npz":
        if key is None:
            raise DiskLoadError(f"missing required key for npz file '{path}'")
        try:
            with np.load(path, allow_pickle=False) as z:
                names = [n[:-4] if n.endswith(".npy") else n for n in z.files]
                if key not in names:
                    raise DiskLoadError(f"key '{key}' not found in .npz '{path}'. "
                                        f"Available keys: {', '.join(names)}")
                arr = _npz_member_mmap(path, key)
                if arr is None:
                    arr = z[key]
        except OSError as e:
            raise DiskLoadError(f"failed to read .npz archive from '{path}': {e}") from None
        return _as_matrix(path, arr)
    if ext == "parquet":
        if key is None:
            raise DiskLoadError(f"missing required key for parquet file '{path}'")
        return _load_parquet(path, key)
    raise DiskLoadError(f"unsupported file extension for '{path}', expected .npy, .npz, "
                        "or .parquet")


def _npz_member_mmap(path: str, key: str):
    """The member ``key`` of an .npz archive as a read-only memory map when
    it is stored without compression and in C order; None otherwise."""
    with zipfile.ZipFile(path) as zf:
        try:
            info = zf.getinfo(key + ".npy")
        except KeyError:
            return None
        if info.compress_type != zipfile.ZIP_STORED:
            return None
        with open(path, "rb") as f:
            # local file header: 30 bytes, then the name and extra field
            f.seek(info.header_offset)
            local = f.read(30)
            if local[:4] != b"PK\x03\x04":
                return None
            name_len = int.from_bytes(local[26:28], "little")
            extra_len = int.from_bytes(local[28:30], "little")
            start = info.header_offset + 30 + name_len + extra_len
            f.seek(start)
            version = np.lib.format.read_magic(f)
            if version == (1, 0):
                shape, fortran, dtype = np.lib.format.read_array_header_1_0(f)
            elif version == (2, 0):
                shape, fortran, dtype = np.lib.format.read_array_header_2_0(f)
            else:
                return None
            data = f.tell()
    if fortran or dtype.hasobject or len(shape) == 0:
        return None
    return np.memmap(path, dtype=dtype, mode="r", offset=data, shape=tuple(shape), order="C")


def _load_parquet(path: str, key: str) -> np.ndarray:
    """load_parquet, src/python_io.rs:167-236 (list / fixed-size-list f32)."""
    import pyarrow as pa
    import pyarrow.parquet as pq

    try:
        table = pq.read_table(path)
    except (OSError, pa.ArrowException) as e:
        raise DiskLoadError(f"failed to read parquet file '{path}': {e}") from None
    if key not in table.column_names:
        raise DiskLoadError(f"column '{key}' was not found in parquet file '{path}'")
    col = table.column(key)
    t = col.type
    if not ((pa.types.is_list(t) or pa.types.is_fixed_size_list(t) or pa.types.is_large_list(t))
            and pa.types.is_float32(t.value_type)):
        raise DiskLoadError(f"parquet key '{key}' in '{path}' is not an array/list<f32> column")
    if col.null_count:
        raise DiskLoadError(f"parquet key '{key}' in '{path}' contains null values, which "
                            "are unsupported")
    rows = []
    dims = None
    for chunk in col.chunks:
        lengths = chunk.value_lengths().to_numpy(zero_copy_only=False) if not \
            pa.types.is_fixed_size_list(t) else np.full(len(chunk), t.list_size)
        flat = chunk.flatten()
        if flat.null_count:
            raise DiskLoadError(f"parquet key '{key}' in '{path}' contains null values, which "
                                "are unsupported")
        vals = flat.to_numpy(zero_copy_only=False).astype(np.float32, copy=False)
        if len(lengths) == 0:
            continue
        d = int(lengths[0])
        if np.any(lengths != d) or (dims is not None and d != dims):
            raise DiskLoadError(f"parquet key '{key}' in '{path}' has inconsistent row "
                                "dimensions")
        dims = d
        rows.append(vals.reshape(len(lengths), d))
    if not rows or dims == 0:
        raise DiskLoadError(f"matrix loaded from '{path}' is empty or has zero dimensions")
    return np.ascontiguousarray(np.concatenate(rows, axis=0))
