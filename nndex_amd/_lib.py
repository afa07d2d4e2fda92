"""ctypes binding of libnndex_hip.so (the C ABI declared in include/nndex_hip.h).

The library is built in-tree by ``python -m nndex_amd.build``.  Loading fails
loudly when it is missing: there is no CPU fallback for any entry point.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnndex_hip.so")

# nndex_status values (include/nndex_hip.h)
OK = 0
SHAPE_MISMATCH = 1
QUERY_DIM_MISMATCH = 2
INVALID_TOP_K = 3
NON_FINITE_INPUT = 4
ZERO_NORM = 5
CPU_BACKEND_UNAVAILABLE = 6
GPU_BACKEND_UNAVAILABLE = 7
GPU_ADAPTER_NOT_FOUND = 10
GPU_REQUEST_DEVICE = 11
GPU_MAP_READ_FAILED = 12
GPU_DEVICE_INDEX_OUT_OF_RANGE = 13
INVALID_ARGUMENT = 20
OUT_OF_MEMORY = 21
HIP_ERROR = 22
UNSUPPORTED = 23

STATUS_NAMES = {
    SHAPE_MISMATCH: "ShapeMismatch",
    QUERY_DIM_MISMATCH: "QueryDimensionalityMismatch",
    INVALID_TOP_K: "InvalidTopK",
    NON_FINITE_INPUT: "NonFiniteInput",
    ZERO_NORM: "ZeroNorm",
    CPU_BACKEND_UNAVAILABLE: "CpuBackendUnavailable",
    GPU_BACKEND_UNAVAILABLE: "GpuBackendUnavailable",
    GPU_ADAPTER_NOT_FOUND: "AdapterNotFound",
    GPU_REQUEST_DEVICE: "RequestDevice",
    GPU_MAP_READ_FAILED: "MapReadFailed",
    GPU_DEVICE_INDEX_OUT_OF_RANGE: "DeviceIndexOutOfRange",
    INVALID_ARGUMENT: "InvalidArgument",
    OUT_OF_MEMORY: "OutOfMemory",
    HIP_ERROR: "HipError",
    UNSUPPORTED: "Unsupported",
}


class NNdexStatusError(RuntimeError):
    """A non-OK nndex_status with the library's message."""

    def __init__(self, code: int, message: str):
        self.code = code
        self.kind = STATUS_NAMES.get(code, str(code))
        super().__init__(message or self.kind)


class IndexOptions(ctypes.Structure):
    """nndex_index_options == IndexOptions (src/lib.rs:63-89)."""

    _fields_ = [
        ("normalized", ctypes.c_int32),
        ("approx", ctypes.c_int32),
        ("backend", ctypes.c_int32),
        ("enable_cache", ctypes.c_int32),
        ("gpu_device_index", ctypes.c_int64),
    ]


class GpuDeviceInfo(ctypes.Structure):
    """nndex_gpu_device_info == GpuDeviceInfo (src/gpu.rs:974-988)."""

    _fields_ = [
        ("index", ctypes.c_uint64),
        ("name", ctypes.c_char * 128),
        ("driver", ctypes.c_char * 64),
        ("driver_info", ctypes.c_char * 128),
        ("backend", ctypes.c_char * 32),
        ("device_type", ctypes.c_char * 32),
    ]

    def to_dict(self) -> dict:
        return {
            "index": int(self.index),
            "name": self.name.decode(),
            "driver": self.driver.decode(),
            "driver_info": self.driver_info.decode(),
            "backend": self.backend.decode(),
            "device_type": self.device_type.decode(),
        }


class IvfInfo(ctypes.Structure):
    """nndex_ivf_info: summary of a built IVF index."""

    _fields_ = [
        ("rows", ctypes.c_uint64),
        ("padded_dims", ctypes.c_uint64),
        ("nlist", ctypes.c_uint64),
        ("iters", ctypes.c_uint64),
        ("rechecked", ctypes.c_uint64),
        ("build_ms", ctypes.c_double),
    ]


class ServerOptions(ctypes.Structure):
    """nndex_server_options."""

    _fields_ = [
        ("slots", ctypes.c_uint32),
        ("max_batch_queries", ctypes.c_uint32),
        ("max_k", ctypes.c_uint32),
        ("ivf_nprobe_max", ctypes.c_uint32),
        ("hbm_budget_bytes", ctypes.c_uint64),
        ("use_graphs", ctypes.c_uint32),
    ]


class ServerStats(ctypes.Structure):
    """nndex_server_stats."""

    _fields_ = [
        ("slots", ctypes.c_uint64),
        ("slot_bytes", ctypes.c_uint64),
        ("requests", ctypes.c_uint64),
        ("queries", ctypes.c_uint64),
        ("batches", ctypes.c_uint64),
        ("mean_batch_queries", ctypes.c_double),
        ("p50_ms", ctypes.c_double),
        ("p99_ms", ctypes.c_double),
        ("max_ms", ctypes.c_double),
        ("span_s", ctypes.c_double),
        ("queries_per_s", ctypes.c_double),
        ("requests_per_s", ctypes.c_double),
        ("graph_batches", ctypes.c_uint64),
    ]

    def to_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_P = ctypes.c_void_p
_SZ = ctypes.c_size_t
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_ST = ctypes.c_int
_FP = ctypes.POINTER(ctypes.c_float)
_U64P = ctypes.POINTER(ctypes.c_uint64)
_SZP = ctypes.POINTER(ctypes.c_size_t)

# name -> (restype, argtypes); every symbol include/nndex_hip.h declares.
SIGNATURES = {
    "nndex_last_error": (ctypes.c_char_p, []),
    "nndex_version": (ctypes.c_char_p, []),
    "nndex_gpu_index_new": (_ST, [_FP, _SZ, _SZ, _I32, _I32, _I64, ctypes.POINTER(_P)]),
    "nndex_gpu_index_free": (None, [_P]),
    "nndex_gpu_index_search": (_ST, [_P, _FP, _SZ, _I32, _U64P, _FP, _SZP]),
    "nndex_gpu_index_search_batch": (_ST, [_P, _FP, _SZ, _SZ, _I32, _U64P, _FP, _SZP]),
    "nndex_gpu_index_adapter_info": (_ST, [_P, ctypes.POINTER(GpuDeviceInfo)]),
    "nndex_list_gpu_devices": (_ST, [ctypes.POINTER(GpuDeviceInfo), _SZ, _SZP]),
    "nndex_gpu_index_rows": (_SZ, [_P]),
    "nndex_gpu_index_padded_dims": (_SZ, [_P]),
    "nndex_workspace_new": (_ST, [_P, _SZ, _SZ, ctypes.POINTER(_P)]),
    "nndex_workspace_free": (None, [_P]),
    "nndex_gpu_index_search_device": (_ST, [_P, _P, _P, _SZ, _SZ, _P, _P, _P]),
    "nndex_gpu_index_search_device_noalloc": (_ST, [_P, _P, _SZ, _SZ, _P, _P, _P]),
    "nndex_gpu_index_search_batch_device": (_ST, [_P, _P, _P, _SZ, _SZ, _I32, _SZ, _P, _P, _P,
                                                  _P]),
    "nndex_gpu_index_autotune": (_ST, [_P, _I32, ctypes.c_char_p, _SZ]),
    "nndex_gpu_index_autotune_batch": (_ST, [_P, _I32, ctypes.c_char_p, _SZ]),
    "nndex_gpu_index_set_batch_tiles": (_ST, [_P, _I32, _I32]),
    "nndex_gpu_index_set_launch": (_ST, [_P, ctypes.c_uint32, ctypes.c_uint32]),
    "nndex_gpu_index_set_ivf_params": (_ST, [_P, _SZ, _SZ, _SZ]),
    "nndex_gpu_index_build_ivf": (_ST, [_P, ctypes.POINTER(IvfInfo)]),
    "nndex_gpu_index_ivf_export": (_ST, [_P, _FP, _U64P, _U64P, _U64P]),
    "nndex_gpu_index_ivf_import": (_ST, [_P, _SZ, _FP, _U64P, _U64P, _U64P,
                                         ctypes.POINTER(IvfInfo)]),
    "nndex_gpu_index_search_ivf_device": (_ST, [_P, _P, _P, _SZ, _SZ, _SZ, _P, _P, _P, _P]),
    "nndex_server_new": (_ST, [_P, ctypes.POINTER(ServerOptions), ctypes.POINTER(_P)]),
    "nndex_server_free": (None, [_P]),
    "nndex_server_submit": (_ST, [_P, _FP, _SZ, _SZ, _I32, _SZ, _U64P]),
    "nndex_server_submit_raw": (_ST, [_P, _FP, _SZ, _SZ, _I32, _SZ, _I32, _SZ, _U64P]),
    "nndex_server_wait": (_ST, [_P, ctypes.c_uint64, _U64P, _FP, _U64P,
                                ctypes.POINTER(ctypes.c_double)]),
    "nndex_server_get_stats": (_ST, [_P, ctypes.POINTER(ServerStats)]),
    "nndex_server_reset_stats": (_ST, [_P]),
    "nndex_gpu_index_capture_batch": (_ST, [_P, _P, _P, _SZ, _SZ, _I32, _SZ, _P, _P, _P, _P,
                                            ctypes.POINTER(_P)]),
    "nndex_graph_launch": (_ST, [_P, _P]),
    "nndex_graph_free": (None, [_P]),
    "nndex_device_copy": (_ST, [_P, _P, _SZ, _P]),
    "nndex_debug_mfma_tile": (_ST, [_I32, _P, _P, ctypes.c_uint32, _P]),
    "nndex_new": (_ST, [_FP, _SZ, _SZ, _SZ, ctypes.POINTER(IndexOptions), ctypes.POINTER(_P)]),
    "nndex_free": (None, [_P]),
    "nndex_backend_of": (_I32, [_P]),
    "nndex_rows": (_SZ, [_P]),
    "nndex_dims": (_SZ, [_P]),
    "nndex_gpu_info": (_ST, [_P, ctypes.POINTER(GpuDeviceInfo), ctypes.POINTER(_I32)]),
    "nndex_search": (_ST, [_P, _FP, _SZ, _SZ, _U64P, _FP, _SZP]),
    "nndex_search_batch": (_ST, [_P, _FP, _SZ, _SZ, _SZ, _U64P, _FP, _SZP]),
    "nndex_backend_handle": (_P, [_P]),
    "nndex_device_count": (_ST, [ctypes.POINTER(_I32)]),
    "nndex_set_device": (_ST, [_I32]),
    "nndex_device_malloc": (_ST, [ctypes.POINTER(_P), _SZ]),
    "nndex_device_free": (_ST, [_P]),
    "nndex_memcpy_h2d": (_ST, [_P, _P, _SZ]),
    "nndex_memcpy_d2h": (_ST, [_P, _P, _SZ]),
    "nndex_stream_create": (_ST, [ctypes.POINTER(_P)]),
    "nndex_stream_destroy": (_ST, [_P]),
    "nndex_stream_synchronize": (_ST, [_P]),
    "nndex_event_create": (_ST, [ctypes.POINTER(_P)]),
    "nndex_event_destroy": (_ST, [_P]),
    "nndex_event_record": (_ST, [_P, _P]),
    "nndex_event_synchronize": (_ST, [_P]),
    "nndex_event_elapsed_ms": (_ST, [_P, _P, ctypes.POINTER(ctypes.c_float)]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Load libnndex_hip.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m nndex_amd.build` "
                "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != OK:
        msg = lib().nndex_last_error()
        raise NNdexStatusError(status, msg.decode() if msg else "")
