"""The exact search entry point that allocates nothing
(nndex_gpu_index_search_device_noalloc), called through the C ABI and
checked bit for bit against nndex_gpu_index_search_device on the same index,
on two calls in a row (the second with fewer queries and a smaller k), and
against the oracle's restatement of the reference GPU path's exact search."""

import ctypes

import numpy as np
import pytest

from nndex_amd import _lib

from gpu_util import BackendIndex, DevBuf, DeviceSearch, gpu_available

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the HIP path has no CPU fallback")


def _noalloc(b, dq, nq, k, didx, dsim, stream):
    return _lib.lib().nndex_gpu_index_search_device_noalloc(b.h, dq.ptr, nq, k, didx.ptr,
                                                            dsim.ptr, stream)


def _both(b, ds, qs, k):
    nq = qs.shape[0]
    dq = DevBuf.from_array(qs)
    i0, s0 = DevBuf(nq * k * 8), DevBuf(nq * k * 4)
    i1, s1 = DevBuf(nq * k * 8), DevBuf(nq * k * 4)
    ds.run(dq, nq, k, i0, s0)
    _lib.check(_noalloc(b, dq, nq, k, i1, s1, ds.stream))
    _lib.check(_lib.lib().nndex_stream_synchronize(ds.stream))
    ref = (i0.to_array((nq, k), np.uint64), s0.to_array((nq, k), np.float32))
    got = (i1.to_array((nq, k), np.uint64), s1.to_array((nq, k), np.float32))
    return ref, got


@pytest.mark.parametrize("rows,dims,dup", [(70_000, 100, False), (9_000, 24, True)])
def test_noalloc_equals_the_workspace_entry_point(orc, rows, dims, dup):
    print(f"[case] noalloc rows={rows} dims={dims} dup={dup}", flush=True)
    rng = np.random.default_rng(rows + dims)
    if dup:  # duplicated rows: ties across the k-th place (the reference heap's rule)
        pool = rng.normal(size=(300, dims)).astype(np.float32)
        raw = pool[rng.integers(0, 300, size=rows)]
    else:
        raw = rng.normal(size=(rows, dims)).astype(np.float32)
    m = orc.normalize_rows(raw, rows, dims)
    qs = orc.normalize_rows(rng.normal(size=(9, dims)).astype(np.float32), 9, dims)
    b = BackendIndex(m)
    ds = DeviceSearch(b.h, 16, 64)
    # first call: 9 queries (the workspace entry point takes its MFMA path),
    # second: 3 queries at a smaller k
    for nq, k in ((9, 40), (3, 7)):
        (ri, rs), (gi, gs) = _both(b, ds, qs[:nq], k)
        np.testing.assert_array_equal(gi, ri)
        assert np.array_equal(gs.view(np.uint32), rs.view(np.uint32))
        for q in range(nq):
            oi, osim = orc.gpu_exact_search(m, qs[q], k)
            np.testing.assert_array_equal(gi[q], oi)
            assert np.array_equal(gs[q].view(np.uint32), osim.view(np.uint32))


def test_noalloc_rejects_what_it_cannot_hold(orc):
    print("[case] noalloc_rejects", flush=True)
    rng = np.random.default_rng(3)
    m = orc.normalize_rows(rng.normal(size=(5000, 16)).astype(np.float32), 5000, 16)
    b = BackendIndex(m)
    dq = DevBuf.from_array(m[:2])
    di, dsim = DevBuf(2 * 2000 * 8), DevBuf(2 * 2000 * 4)
    assert _noalloc(b, dq, 2, 0, di, dsim, None) == _lib.INVALID_TOP_K
    # k above 992 needs the large-k selection's buffers (a workspace)
    assert _noalloc(b, dq, 2, 993, di, dsim, None) == _lib.UNSUPPORTED
    assert b"fused scan" in _lib.lib().nndex_last_error()
    assert _noalloc(b, dq, 2, 1985, di, dsim, None) == _lib.UNSUPPORTED
    assert _noalloc(b, dq, 2, 5001, di, dsim, None) == _lib.INVALID_ARGUMENT
    # zero queries: nothing to do
    assert _noalloc(b, dq, 0, 5, di, dsim, None) == _lib.OK
