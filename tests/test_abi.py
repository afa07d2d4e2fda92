"""C-ABI checks that need no GPU: the library loads, exports every symbol the
headers declare, and the host-side front-end raises the reference's errors."""

import ctypes
import os
import re

import numpy as np
import pytest

from nndex_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(nndex_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _declared_symbols()
    assert len(declared) >= 30
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding covers the whole declared surface
    assert declared <= set(_lib.SIGNATURES), sorted(declared - set(_lib.SIGNATURES))


def test_version_and_empty_error():
    L = _lib.lib()
    assert b"gfx950" in L.nndex_version()
    assert L.nndex_last_error() is not None


def _new(values, rows, dims, **kw):
    opts = _lib.IndexOptions(normalized=kw.get("normalized", 0), approx=0,
                             backend=kw.get("backend", 0), enable_cache=1, gpu_device_index=-1)
    h = ctypes.c_void_p()
    v = np.ascontiguousarray(values, np.float32).reshape(-1)
    st = _lib.lib().nndex_new(v.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), v.size, rows,
                              dims, ctypes.byref(opts), ctypes.byref(h))
    return st, _lib.lib().nndex_last_error().decode()


def test_shape_errors_precede_backend(  ):
    # src/lib.rs:1131-1141: rows*dims != len
    st, msg = _new(np.ones(10), 3, 4)
    assert st == _lib.SHAPE_MISMATCH
    assert msg == "matrix shape mismatch: rows=3, dims=4, values=10"
    st, _ = _new(np.ones(0), 0, 8)
    assert st == _lib.SHAPE_MISMATCH


def test_cpu_backend_unavailable_after_validation():
    # GPU-only build, like the reference without its `cpu` feature.
    st, msg = _new(np.eye(8), 8, 8)
    assert st == _lib.CPU_BACKEND_UNAVAILABLE
    assert msg == "cpu backend is unavailable (crate built without `cpu` feature)"
    # normalisation errors come first (src/lib.rs:239-241 before :244)
    bad = np.ones((1, 8), np.float32)
    bad[0, 1] = np.nan
    assert _new(bad, 1, 8)[0] == _lib.NON_FINITE_INPUT
    assert _new(np.zeros((1, 8)), 1, 8)[0] == _lib.ZERO_NORM
    assert _new(np.zeros((1, 8)), 1, 8, normalized=1)[0] == _lib.CPU_BACKEND_UNAVAILABLE


def test_python_surface_errors():
    import nndex_amd

    with pytest.raises(ValueError, match="unsupported backend 'tpu'"):
        nndex_amd.NNdex(np.eye(8, dtype=np.float32), backend="TPU")
    with pytest.raises(ValueError, match="cpu backend is unavailable"):
        nndex_amd.NNdex(np.eye(8, dtype=np.float32))
    with pytest.raises(ValueError, match="consistent row lengths"):
        nndex_amd.NNdex([[1.0, 2.0], [1.0]], backend="gpu")


def test_list_gpu_devices_without_gpu_is_safe():
    import nndex_amd

    devs = nndex_amd.list_gpu_devices()
    assert isinstance(devs, list)
    for d in devs:
        assert d["backend"] and d["device_type"]


def test_io_loaders(tmp_path):
    from nndex_amd.io import DiskLoadError, load_matrix_from_disk

    m = np.arange(24, dtype=np.float32).reshape(3, 8)
    np.save(tmp_path / "m.npy", m)
    assert np.array_equal(load_matrix_from_disk(str(tmp_path / "m.npy")), m)
    np.savez(tmp_path / "m.npz", emb=m)
    assert np.array_equal(load_matrix_from_disk(str(tmp_path / "m.npz"), "emb"), m)
    with pytest.raises(DiskLoadError, match="missing required key for npz"):
        load_matrix_from_disk(str(tmp_path / "m.npz"))
    with pytest.raises(DiskLoadError, match="Available keys: emb"):
        load_matrix_from_disk(str(tmp_path / "m.npz"), "nope")
    with pytest.raises(DiskLoadError, match="unsupported file extension"):
        load_matrix_from_disk(str(tmp_path / "m.csv"))
    pa = pytest.importorskip("pyarrow")
    import pyarrow.parquet as pq

    pq.write_table(pa.table({"v": pa.array(list(m), type=pa.list_(pa.float32()))}),
                   tmp_path / "m.parquet")
    assert np.array_equal(load_matrix_from_disk(str(tmp_path / "m.parquet"), "v"), m)
    with pytest.raises(DiskLoadError, match="was not found"):
        load_matrix_from_disk(str(tmp_path / "m.parquet"), "w")


def test_server_boundary_checks_without_gpu():
    """Argument checks of the serving ABI run before any device call."""
    import ctypes

    from nndex_amd import _lib

    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.nndex_server_new(None, None, ctypes.byref(h)) == 20
    assert b"null pointer" in L.nndex_last_error()
    t = ctypes.c_uint64()
    assert L.nndex_server_submit(None, None, 1, 1, 0, 0, ctypes.byref(t)) == 20
    assert L.nndex_server_wait(None, 1, None, None, None, None) == 20
    L.nndex_server_free(None)
