"""Exact search for k beyond the fused scan's LDS lists (k > 992, up to
k = rows): every row's key, a radix select of the top min(2k - 1, rows), a
sort and the reference heap's tie rule, on the device.

The reference accepts any k <= rows (src/lib.rs:347) and pushes every row in
row order into one bounded heap (src/gpu.rs:575-589, src/topk.rs:44-152);
results are compared bit for bit with the oracle's restatement of that heap
(orc.gpu_exact_search), through the backend boundary's host and device entry
points."""

import numpy as np
import pytest

from gpu_util import BackendIndex, DevBuf, DeviceSearch, gpu_available

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the HIP path has no CPU fallback")


@pytest.fixture(scope="module")
def big(orc):
    rows, dims = 120_000, 64
    rng = np.random.default_rng(77)
    m = orc.normalize_rows(rng.normal(size=(rows, dims)).astype(np.float32), rows, dims)
    qs = orc.normalize_rows(rng.normal(size=(6, dims)).astype(np.float32), 6, dims)
    return m, qs, BackendIndex(m)


def _same(orc, m, q, idx, sim, k):
    ridx, rsim = orc.gpu_exact_search(m, q, k)
    assert len(idx) == len(ridx) == min(k, m.shape[0])
    np.testing.assert_array_equal(np.asarray(idx, np.uint64), ridx)
    assert np.array_equal(np.asarray(sim, np.float32).view(np.uint32), rsim.view(np.uint32))


@pytest.mark.parametrize("k", [993, 2000, 10_000, 120_000])
def test_large_k_host_entry_points(orc, big, k):
    m, qs, b = big
    print(f"[case] large_k host k={k}", flush=True)
    idx, sim = b.search(qs[0], k)
    _same(orc, m, qs[0], idx, sim, k)
    if k <= 10_000:
        bidx, bsim, lens = b.search_batch(qs[:5], k)  # a group of 4, then 1
        for i in range(5):
            assert int(lens[i]) == k
            _same(orc, m, qs[i], bidx[i], bsim[i], k)


def test_large_k_device_entry_point(orc, big):
    m, qs, b = big
    k = 2500
    print(f"[case] large_k device k={k}", flush=True)
    ds = DeviceSearch(b.h, 8, k)
    dq = DevBuf.from_array(qs)
    di, dsim = DevBuf(6 * k * 8), DevBuf(6 * k * 4)
    ds.run(dq, 6, k, di, dsim)
    idx, sim = di.to_array((6, k), np.uint64), dsim.to_array((6, k), np.float32)
    for i in range(6):
        _same(orc, m, qs[i], idx[i], sim[i], k)
    # a second call with fewer queries reuses the selection buffers
    ds.run(dq, 2, 1200, di, dsim)
    idx, sim = di.to_array((2, 1200), np.uint64), dsim.to_array((2, 1200), np.float32)
    for i in range(2):
        _same(orc, m, qs[i], idx[i], sim[i], 1200)


def test_large_k_ties_follow_the_reference_heap(orc):
    """Duplicated rows (~300 copies each): rows tie bit for bit across the
    k-th place, where the heap's choice differs from score-then-row order."""
    rows, dims = 60_000, 24
    rng = np.random.default_rng(12)
    pool = rng.normal(size=(200, dims)).astype(np.float32)
    pick = rng.integers(0, 200, size=rows)
    m = orc.normalize_rows(pool[pick], rows, dims)
    pn = orc.normalize_rows(pool, 200, dims)
    qs = orc.normalize_rows(rng.normal(size=(4, dims)).astype(np.float32), 4, dims)
    b = BackendIndex(m)
    print("[case] large_k ties", flush=True)
    shifted = 0
    for k in (1000, 1777, 4321):
        bidx, bsim, _ = b.search_batch(qs, k)
        for i in range(4):
            _same(orc, m, qs[i], bidx[i], bsim[i], k)
            # score-then-row order with the reference dot (copies score alike)
            sp = np.array([orc.dot_product(pn[j], qs[i]) for j in range(200)], np.float32)
            sc = sp[pick].astype(np.float64)
            canon = np.lexsort((np.arange(rows), -sc))[:k]
            shifted += not np.array_equal(np.sort(canon), np.sort(bidx[i].astype(np.int64)))
    assert shifted > 0  # the case exercises the heap's own choice
