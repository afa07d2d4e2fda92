"""Parity of the HIP exact-search path against the CPU oracle and FP64.

Bar (north star): top-k indices identical except where FP64 scores tie within
1e-6; FP32 scores within 1e-5 of FP64.  Because the device scan reproduces the
reference's dot-product order, scores are also compared bit-for-bit with the
oracle's restatement of the reference CPU path.
"""

import numpy as np
import pytest

import nndex_amd
from nndex_amd import synth

from gpu_util import (BackendIndex, DevBuf, DeviceSearch, assert_topk_parity, gpu_available)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the HIP path has no CPU fallback")


def _print_case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


# ---------------------------------------------------------------------------
# Golden vector (notebooks/python_bindings_demo.ipynb:102-103) through NNdex
# ---------------------------------------------------------------------------

def test_golden_single_query_through_frontend(orc):
    _print_case("golden_50k_single")
    matrix, query, queries = synth.golden_50k()
    index = nndex_amd.NNdex(matrix, normalized=False, backend="gpu", enable_cache=False)
    assert index.backend == "gpu" and index.rows == 50_000 and index.dims == 128
    idx, sim = index.search(query, k=5)
    assert idx.dtype == np.uint64 and sim.dtype == np.float32
    assert idx.tolist() == synth.GOLDEN_50K_INDICES
    assert [float(s) for s in sim] == synth.GOLDEN_50K_SCORES  # bit-exact


def test_golden_batch_through_frontend(orc):
    _print_case("golden_50k_batch4")
    matrix, _, queries = synth.golden_50k()
    index = nndex_amd.NNdex(matrix, normalized=False, backend="gpu", enable_cache=False)
    bidx, bsim = index.search(queries, k=5)
    assert bidx.shape == (4, 5) and bsim.shape == (4, 5)
    m = orc.normalize_rows(matrix, 50_000, 128)
    qn = orc.normalize_rows(queries, 4, 128)
    for i in range(4):
        oidx, osim, _ = orc.cpu_search(m, qn[i], 5)
        assert bidx[i].tolist() == oidx.tolist()
        assert bsim[i].tolist() == osim.tolist()  # same arithmetic as the oracle
        full = orc.fp64_scores(matrix, queries[i], normalized=False)
        fidx, _ = orc.fp64_topk(matrix, queries[i], 5, normalized=False)
        assert_topk_parity(bidx[i], bsim[i], fidx[0], full)


# ---------------------------------------------------------------------------
# Parity sweep over shapes, k and options
# ---------------------------------------------------------------------------

CASES = [
    # rows, dims, k, normalized
    (1, 8, 1, True),
    (3, 7, 5, False),
    (63, 13, 10, False),
    (64, 16, 64, False),
    (65, 100, 7, False),
    (1000, 384, 10, True),
    (4097, 128, 25, False),
    (20000, 96, 1, False),
    (20000, 96, 100, True),
    (30001, 40, 1000, False),
    (5000, 1024, 17, False),
]


@pytest.mark.parametrize("rows,dims,k,normalized", CASES)
def test_parity_sweep(orc, rows, dims, k, normalized):
    _print_case("parity", rows=rows, dims=dims, k=k, normalized=normalized)
    rng = np.random.default_rng(rows * 31 + dims)
    m = rng.normal(size=(rows, dims)).astype(np.float32)
    qs = rng.normal(size=(3, dims)).astype(np.float32)
    if normalized:
        m /= np.linalg.norm(m, axis=1, keepdims=True)
        qs /= np.linalg.norm(qs, axis=1, keepdims=True)
    index = nndex_amd.NNdex(m, normalized=normalized, backend="gpu", enable_cache=False)
    pd = orc.padded_dims(dims)
    om = orc.pad_rows(m, rows, dims, pd) if normalized else orc.normalize_rows(m, rows, dims)
    for q in qs:
        idx, sim = index.search(q, k=k)
        oq = orc.pad_rows(q, 1, dims, pd)[0] if normalized else orc.normalize_vector(q)
        oidx, osim, _ = orc.cpu_search(om, oq, k)
        assert len(idx) == min(k, rows)
        assert idx.tolist() == oidx.tolist()
        assert sim.tolist() == osim.tolist()
        full = orc.fp64_scores(m, q, normalized=normalized)
        fidx, _ = orc.fp64_topk(m, q, k, normalized=normalized)
        assert_topk_parity(idx, sim, fidx[0], full)
    # batch: queries prepared on the device with normalize_rows / pad_rows
    # semantics, then the same per-query arithmetic as the single path
    bidx, bsim = index.search(qs, k=k)
    obq = orc.pad_rows(qs, 3, dims, pd) if normalized else orc.normalize_rows(qs, 3, dims)
    for i in range(3):
        oidx, osim, _ = orc.cpu_search(om, obq[i], k)
        assert bidx[i].tolist() == oidx.tolist() and bsim[i].tolist() == osim.tolist()


def test_ties_resolve_to_lower_index(orc):
    """Duplicate rows score identically; the total order keeps lower indices
    (src/topk.rs:143-148, topk_from_scores' comparator)."""
    _print_case("ties")
    rng = np.random.default_rng(3)
    base = rng.normal(size=(50, 32)).astype(np.float32)
    m = np.concatenate([base] * 40)  # row r duplicates row r % 50
    index = nndex_amd.NNdex(m, normalized=False, backend="gpu", enable_cache=False)
    q = base[7] + 0.01
    idx, sim = index.search(q, k=40)
    assert sorted(set((idx % 50).tolist())) == [7]
    assert idx.tolist() == [7 + 50 * i for i in range(40)]
    assert len(set(sim.tolist())) == 1


def test_ties_follow_the_reference_heap(orc):
    """Rows tied at the k-th score, with higher-scoring rows placed after them:
    the reference GPU path's heap (rows pushed in order, lowest row evicted
    first among equal minima) keeps later tied rows than the (score desc, row
    asc) order would.  Single queries, batches (the certified MFMA path hands
    such queries to the scan) and the device entry points must all agree with
    the oracle's restatement of that path."""
    _print_case("heap_ties")
    rng = np.random.default_rng(11)
    pool = rng.normal(size=(6, 32)).astype(np.float32)
    rows = np.concatenate([rng.integers(1, 6, size=3000), np.zeros(20, np.int64)])
    m = np.ascontiguousarray(pool[rows])
    qs = np.ascontiguousarray(pool[rng.integers(0, 6, size=12)] +
                              0.2 * rng.normal(size=(12, 32)).astype(np.float32))
    om = orc.normalize_rows(m, m.shape[0], 32)
    oq = orc.normalize_rows(qs, 12, 32)
    index = nndex_amd.NNdex(m, normalized=False, backend="gpu", enable_cache=False)
    shifted = 0
    for k in (2, 25, 60):
        bidx, bsim = index.search(qs, k=k)
        for i in range(12):
            # single queries are normalised as vectors, batches as rows
            sidx, ssim = index.search(qs[i], k=k)
            o1idx, o1sim = orc.gpu_exact_search(om, orc.normalize_vector(qs[i]), k)
            assert sidx.tolist() == o1idx.tolist() and ssim.tolist() == o1sim.tolist(), (k, i)
            oidx, osim = orc.gpu_exact_search(om, oq[i], k)
            assert bidx[i].tolist() == oidx.tolist() and bsim[i].tolist() == osim.tolist(), (k, i)
            sc = [orc.dot_product(om[r], oq[i]) for r in range(om.shape[0])]
            canon = sorted(range(om.shape[0]), key=lambda r: (-sc[r], r))[:k]
            shifted += canon != oidx.tolist()
    assert shifted > 0  # the case exercises the heap's own choice


def test_final_merge_with_every_block_maximum_tied(orc):
    """200,000 rows drawn from 8 vectors: every scan block's best keys tie at
    the top score, so every published key reaches the final merge's bound.
    At k = 10 the 196 blocks x 19 kept keys overflow the merge's LDS pool
    (512 keys for one query, 2,048 for the 4-query form of a 3-query batch)
    and take the fallback in which every wave offers its share of the table;
    at k = 1 the 196 keys fit the pool.  Single queries and the batch must
    equal the oracle's restatement of the reference GPU path (its heap keeps
    the lowest tied rows here)."""
    _print_case("final_merge_ties", rows=200000, dims=32)
    rng = np.random.default_rng(21)
    pool = rng.normal(size=(8, 32)).astype(np.float32)
    m = np.ascontiguousarray(pool[rng.integers(0, 8, size=200000)])
    qs = np.ascontiguousarray(pool[[3, 5, 0]] + 0.05 * rng.normal(size=(3, 32)).astype(np.float32))
    om = orc.normalize_rows(m, m.shape[0], 32)
    oq = orc.normalize_rows(qs, 3, 32)
    index = nndex_amd.NNdex(m, normalized=False, backend="gpu", enable_cache=False)
    for k in (10, 1):
        bidx, bsim = index.search(qs, k=k)
        for i in range(3):
            sidx, ssim = index.search(qs[i], k=k)
            o1idx, o1sim = orc.gpu_exact_search(om, orc.normalize_vector(qs[i]), k)
            assert sidx.tolist() == o1idx.tolist() and ssim.tolist() == o1sim.tolist(), (k, i)
            oidx, osim = orc.gpu_exact_search(om, oq[i], k)
            assert bidx[i].tolist() == oidx.tolist() and bsim[i].tolist() == osim.tolist(), (k, i)
            assert len(set(oidx.tolist())) == k and len(set(osim.tolist())) == 1


def test_all_negative_scores_and_k_capping():
    """src/lib.rs:1466-1506 and :1374-1425."""
    _print_case("negatives")
    m = np.zeros((5, 8), np.float32)
    m[:, 0] = 1.0
    index = nndex_amd.NNdex(m, normalized=True, backend="gpu", enable_cache=False)
    q = np.zeros((4, 8), np.float32)
    q[:, 0] = -1.0
    idx, sim = index.search(q, k=3)
    assert idx.shape == (4, 3) and np.all(sim == -1.0) and np.all(idx < 5)
    idx, sim = index.search(q[0], k=100)
    assert len(idx) == 5 and idx.tolist() == [0, 1, 2, 3, 4]


def test_reference_kats_through_gpu():
    """src/lib.rs:824-889 and tests_python/test_bindings.py:9-33."""
    _print_case("kats")
    m = np.array([[1, 0, 0, 0, 0, 0, 0, 0], [0, 1, 0, 0, 0, 0, 0, 0], [1, 1, 0, 0, 0, 0, 0, 0]],
                 np.float32)
    index = nndex_amd.NNdex(m, normalized=False, backend="gpu", enable_cache=False)
    idx, sim = index.search([0.8, 0.6, 0, 0, 0, 0, 0, 0], k=2)
    assert idx.tolist() == [2, 0]
    assert abs(sim[0] - 0.98994946) < 1e-5 and abs(sim[1] - 0.8) < 1e-6
    eye = nndex_amd.NNdex(np.eye(8, dtype=np.float32), normalized=True, backend="gpu",
                          enable_cache=False)
    bidx, bsim = eye.search(np.eye(8, dtype=np.float32)[:2], k=2)
    assert bidx[:, 0].tolist() == [0, 1]
    one = nndex_amd.NNdex(np.ones((3, 7), np.float32), backend="gpu", enable_cache=False)
    assert one.dims == 7 and len(one.search(np.ones(7, np.float32), k=1)[0]) == 1


def test_error_paths_on_gpu():
    _print_case("errors")
    bad = np.ones((4, 8), np.float32)
    bad[2, 3] = np.nan
    with pytest.raises(ValueError, match="all input values must be finite"):
        nndex_amd.NNdex(bad, backend="gpu", enable_cache=False)
    bad[2, 3] = np.inf
    with pytest.raises(ValueError, match="all input values must be finite"):
        nndex_amd.NNdex(bad, normalized=True, backend="gpu", enable_cache=False)
    z = np.ones((4, 8), np.float32)
    z[1] = 0
    with pytest.raises(ValueError, match="zero-norm"):
        nndex_amd.NNdex(z, backend="gpu", enable_cache=False)
    index = nndex_amd.NNdex(np.eye(8, dtype=np.float32), backend="gpu", enable_cache=False)
    with pytest.raises(RuntimeError, match="top-k must be greater than zero"):
        index.search(np.ones(8, np.float32), k=0)
    with pytest.raises(ValueError, match="query dims mismatch: expected 8, got 4"):
        index.search(np.ones(4, np.float32), k=1)
    with pytest.raises(RuntimeError, match="all input values must be finite"):
        index.search(np.array([np.nan] + [0] * 7, np.float32), k=1)
    with pytest.raises(RuntimeError, match="zero-norm"):
        index.search(np.zeros(8, np.float32), k=1)
    with pytest.raises(RuntimeError, match="zero-norm"):
        index.search(np.zeros((2, 8), np.float32), k=1)
    with pytest.raises(ValueError, match="gpu device index 9999 is out of range"):
        nndex_amd.NNdex(np.eye(8, dtype=np.float32), backend="gpu", gpu_device_index=9999)


def test_gpu_info_and_devices():
    _print_case("info")
    devs = nndex_amd.list_gpu_devices()
    assert len(devs) >= 1 and devs[0]["index"] == 0
    index = nndex_amd.NNdex(np.eye(8, dtype=np.float32), backend="gpu", enable_cache=False)
    info = index.gpu_info
    assert info is not None and info["index"] == 0 and info["name"]


def test_constructor_cache_reuses_backend():
    """src/lib.rs:1025-1074: same source pointer and shape -> same backend."""
    _print_case("ctor_cache")
    m = np.ones((64, 16), np.float32)
    a = nndex_amd.NNdex(m, backend="gpu", enable_cache=True)
    b = nndex_amd.NNdex(m, backend="gpu", enable_cache=True)
    q = np.ones(16, np.float32)
    assert a.search(q, 5)[0].tolist() == b.search(q, 5)[0].tolist()


# ---------------------------------------------------------------------------
# Backend boundary and device-resident entry point
# ---------------------------------------------------------------------------

def test_backend_boundary_and_device_path(orc):
    _print_case("device_path")
    rng = np.random.default_rng(11)
    rows, dims = 70_001, 64
    m = orc.normalize_rows(rng.normal(size=(rows, dims)).astype(np.float32), rows, dims)
    qs = orc.normalize_rows(rng.normal(size=(11, dims)).astype(np.float32), 11, dims)
    b = BackendIndex(m)
    ds = DeviceSearch(b.h, 16, 64)
    dq = DevBuf.from_array(qs)
    for k in (1, 10, 64):
        didx, dsim = DevBuf(11 * k * 8), DevBuf(11 * k * 4)
        for _rep in range(3):  # the ticket and thresholds reset between launches
            ds.run(dq, 11, k, didx, dsim)
            gi = didx.to_array((11, k), np.uint64)
            gs = dsim.to_array((11, k), np.float32)
            for i in range(11):
                oidx, osim, _ = orc.cpu_search(m, qs[i], k)
                assert gi[i].tolist() == oidx.tolist() and gs[i].tolist() == osim.tolist()
    hidx, hsim = b.search(qs[0], 10)
    bidx, bsim, lens = b.search_batch(qs, 10)
    assert lens.tolist() == [10] * 11 and bidx[0].tolist() == hidx.tolist()


def test_batch_position_independence(orc):
    """A query's result does not depend on its position or its batch mates."""
    _print_case("position_independence")
    rng = np.random.default_rng(12)
    rows, dims = 40_000, 128
    m = orc.normalize_rows(rng.normal(size=(rows, dims)).astype(np.float32), rows, dims)
    qs = orc.normalize_rows(rng.normal(size=(13, dims)).astype(np.float32), 13, dims)
    b = BackendIndex(m)
    i1, s1, _ = b.search_batch(qs, 10)
    perm = rng.permutation(13)
    i2, s2, _ = b.search_batch(qs[perm], 10)
    assert np.array_equal(i1[perm], i2) and np.array_equal(s1[perm], s2)
    for j in range(13):
        si, ss = b.search(qs[j], 10)
        assert si.tolist() == i1[j].tolist() and ss.tolist() == s1[j].tolist()


# ---------------------------------------------------------------------------
# Full size: 1,000,000 x 384 Gaussian mixture, k=10
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def mixture_1m():
    return synth.mixture_workload(n_queries=64)


def test_full_size_1m_x_384(orc, mixture_1m):
    _print_case("full_1m_384")
    base, queries = mixture_1m
    b = BackendIndex(base)
    ds = DeviceSearch(b.h, 64, 10)
    dq = DevBuf.from_array(queries)
    didx, dsim = DevBuf(64 * 10 * 8), DevBuf(64 * 10 * 4)
    ds.run(dq, 64, 10, didx, dsim)
    gi = didx.to_array((64, 10), np.uint64)
    gs = dsim.to_array((64, 10), np.float32)
    for i in range(6):  # oracle + FP64 on a sample of the queries
        oidx, osim, strategy = orc.cpu_search(base, queries[i], 10)
        assert strategy == "ParallelFold"
        assert gi[i].tolist() == oidx.tolist() and gs[i].tolist() == osim.tolist()
        full = orc.fp64_scores(base, queries[i], normalized=True)
        fidx, _ = orc.fp64_topk(base, queries[i], 10, normalized=True)
        assert_topk_parity(gi[i], gs[i], fidx[0], full)
    # size-independent properties for every query: sorted, unique, in range
    assert np.all(np.diff(gs, axis=1) <= 0)
    assert all(len(set(r.tolist())) == 10 for r in gi)
    assert gi.max() < base.shape[0]
    # each query's reported scores equal a direct recomputation of those rows
    for i in range(64):
        rows = base[gi[i].astype(np.int64)]
        direct = [orc.dot_product(r, queries[i]) for r in rows]
        assert direct == gs[i].tolist()


def test_frontend_autotune_keeps_results(orc):
    """NNdex.autotune times the scan and batch tile shapes on the index's own
    data; searches before and after return the same results."""
    _print_case("frontend_autotune")
    rng = np.random.default_rng(21)
    m = rng.normal(size=(30_000, 96)).astype(np.float32)
    qs = rng.normal(size=(40, 96)).astype(np.float32)
    index = nndex_amd.NNdex(m, normalized=False, backend="gpu", enable_cache=False)
    before = index.search(qs, k=10), index.search(qs[0], k=10)
    report = index.autotune(reps=2)
    assert "chosen flags=" in report and report.count("chosen ") >= 3
    after = index.search(qs, k=10), index.search(qs[0], k=10)
    for (bi, bs), (ai, as_) in zip(before, after):
        assert np.array_equal(bi, ai) and np.array_equal(bs, as_)


# ---------------------------------------------------------------------------
# Committed fixtures (tests/golden/fixtures.npz)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("name", ["dims37", "ties", "negative", "k_gt_rows"])
def test_committed_fixtures_through_frontend(name):
    """Single queries, a batch of 6 (fused scan) and the same queries twice
    (12: the certified MFMA path) equal the committed reference-path results
    bit for bit; scores agree with the committed FP64 values to 1e-5."""
    import os

    _print_case(f"fixture_{name}")
    path = os.path.join(os.path.dirname(__file__), "golden", "fixtures.npz")
    with np.load(path) as f:
        fx = {key: f[key] for key in f.files if key.startswith(name + "/")}
    matrix, queries = fx[f"{name}/matrix"], fx[f"{name}/queries"]
    k = int(fx[f"{name}/k"])
    index = nndex_amd.NNdex(matrix, normalized=False, backend="gpu", enable_cache=False)
    for i in range(queries.shape[0]):
        idx, sim = index.search(queries[i], k=k)
        assert idx.tolist() == fx[f"{name}/single_idx"][i].tolist(), i
        assert np.array_equal(sim.view(np.uint32), fx[f"{name}/single_sim"][i].view(np.uint32)), i
        np.testing.assert_allclose(sim, fx[f"{name}/fp64_sim"][i], atol=1e-5, rtol=0)
    for qs in (queries, np.vstack([queries, queries])):
        bidx, bsim = index.search(qs, k=k)
        for i in range(qs.shape[0]):
            j = i % queries.shape[0]
            assert bidx[i].tolist() == fx[f"{name}/batch_idx"][j].tolist(), (qs.shape, i)
            assert np.array_equal(bsim[i].view(np.uint32),
                                  fx[f"{name}/batch_sim"][j].view(np.uint32)), (qs.shape, i)
