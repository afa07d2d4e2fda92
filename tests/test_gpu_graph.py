"""HIP graph capture of the batched search: every replay, with eager work and
synchronisations between replays and new queries written into the captured
buffer, equals a fresh eager search of the same queries."""

import ctypes

import numpy as np
import pytest

from nndex_amd import _lib, synth

from gpu_util import BackendIndex, DevBuf, DeviceSearch, check, gpu_available

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the HIP path has no CPU fallback")


def _eager(b, ds, q, k, prec):
    nq = q.shape[0]
    dq = DevBuf.from_array(q)
    di, dsim, dfl = DevBuf(nq * k * 8), DevBuf(nq * k * 4), DevBuf(nq)
    check(_lib.lib().nndex_gpu_index_search_batch_device(b.h, ds.ws, dq.ptr, nq, k, prec, 0,
                                                         di.ptr, dsim.ptr, dfl.ptr, ds.stream))
    check(_lib.lib().nndex_stream_synchronize(ds.stream))
    return (di.to_array((nq, k), np.uint64), dsim.to_array((nq, k), np.float32),
            dfl.to_array((nq,), np.uint8))


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_graph_replays_equal_eager(prec):
    print(f"[case] graph prec={prec}", flush=True)
    L = _lib.lib()
    base, bank = synth.mixture_workload(rows=80_000, dims=128, n_queries=1024)
    b = BackendIndex(base)
    nq, k = 200, 10
    cap = DeviceSearch(b.h, nq, 128)   # workspace owned by the graph
    eager = DeviceSearch(b.h, nq, 128)  # a second workspace for eager runs
    dq = DevBuf.from_array(bank[:nq])
    di, dsim, dfl = DevBuf(nq * k * 8), DevBuf(nq * k * 4), DevBuf(nq)
    g = ctypes.c_void_p()
    check(L.nndex_gpu_index_capture_batch(b.h, cap.ws, dq.ptr, nq, k, prec, 0, di.ptr, dsim.ptr,
                                          dfl.ptr, cap.stream, ctypes.byref(g)))
    try:
        for rep in range(4):
            q = bank[(rep * 211) % 800:(rep * 211) % 800 + nq]
            check(L.nndex_memcpy_h2d(dq.ptr, np.ascontiguousarray(q).ctypes.data, q.nbytes))
            # eager work on another stream between replays
            ref = _eager(b, eager, q, k, prec)
            check(L.nndex_graph_launch(g, cap.stream))
            check(L.nndex_stream_synchronize(cap.stream))
            got = (di.to_array((nq, k), np.uint64), dsim.to_array((nq, k), np.float32),
                   dfl.to_array((nq,), np.uint8))
            assert np.array_equal(got[0], ref[0]), rep
            assert np.array_equal(got[1].view(np.uint32), ref[1].view(np.uint32)), rep
            assert np.array_equal(got[2], ref[2]), rep
    finally:
        L.nndex_graph_free(g)
