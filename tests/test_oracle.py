"""Pins the CPU oracle (oracle/) before it is trusted as the parity checker.

Each check names the reference test or artefact it restates (fblissjr/nndex
v0.2.1).  The golden vector of the 50,000 x 128 case pins the dot-product and
normalisation arithmetic bit-for-bit; closed forms pin the FP64 oracle; the
reference's own KATs pin top-k, strategy selection and IVF bookkeeping.
"""

import numpy as np
import pytest

from nndex_amd import synth


# ---------------------------------------------------------------------------
# Golden vector: notebooks/python_bindings_demo.ipynb:33-37, 102-103
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def golden():
    return synth.golden_50k()


def test_golden_single_query_bit_exact(orc, golden):
    matrix, query, _ = golden
    m = orc.normalize_rows(matrix, *matrix.shape)
    q = orc.normalize_vector(query)
    idx, sim, strategy = orc.cpu_search(m, q, 5)
    assert strategy == "ParallelChunked" or strategy == "ParallelFold"
    assert idx.tolist() == synth.GOLDEN_50K_INDICES
    # float32 scores are compared bit-for-bit with the published values
    assert [float(s) for s in sim] == synth.GOLDEN_50K_SCORES


def test_golden_scores_within_1e5_of_fp64(orc, golden):
    matrix, query, queries = golden
    fidx, fsim = orc.fp64_topk(matrix, query, 5, normalized=False)
    assert fidx[0].tolist() == synth.GOLDEN_50K_INDICES
    np.testing.assert_allclose(fsim[0], synth.GOLDEN_50K_SCORES, atol=1e-5, rtol=0)


def test_golden_batch_matches_fp64_and_single(orc, golden):
    matrix, _, queries = golden
    m = orc.normalize_rows(matrix, *matrix.shape)
    qn = orc.normalize_rows(queries, *queries.shape)
    bidx, bsim, lens, strategy = orc.cpu_search_batch(m, qn, 5)
    assert strategy == "Gemm"  # profiles.rs:151-185 at dims=128, Q=4
    assert lens.tolist() == [5] * 4
    fidx, fsim = orc.fp64_topk(matrix, queries, 5, normalized=False)
    assert np.array_equal(bidx, fidx)
    np.testing.assert_allclose(bsim, fsim, atol=1e-5, rtol=0)
    for i in range(4):
        sidx, ssim, _ = orc.cpu_search(m, qn[i], 5)
        assert sidx.tolist() == bidx[i].tolist()
        np.testing.assert_allclose(ssim, bsim[i], atol=1e-5, rtol=0)


def test_wrong_dot_order_misses_golden(orc, golden):
    """A plausible mistake (eight f32 accumulators instead of simsimd's four
    FMA lanes) must change a golden score; the check is not vacuous."""
    matrix, query, _ = golden
    m = orc.normalize_rows(matrix, *matrix.shape)
    q = orc.normalize_vector(query)
    alt = [orc.dot_ndarray(m[i], q) for i in synth.GOLDEN_50K_INDICES]
    assert alt != synth.GOLDEN_50K_SCORES


# ---------------------------------------------------------------------------
# FP64 oracle: closed forms
# ---------------------------------------------------------------------------

def test_fp64_closed_form_scaled_basis(orc):
    dims = 12
    scales = np.arange(1, dims + 1, dtype=np.float32)
    matrix = np.diag(scales).astype(np.float32)  # row i = i-th basis vector scaled
    q = np.array([3, 4] + [0] * (dims - 2), np.float32)
    idx, sim = orc.fp64_topk(matrix, q, 3, normalized=False)
    assert idx[0].tolist() == [1, 0, 2]  # 4/5 > 3/5 > 0 (ties at 0: lower index)
    np.testing.assert_allclose(sim[0], [0.8, 0.6, 0.0], atol=1e-15, rtol=0)


def test_fp64_closed_form_negatives_and_normalized_flag(orc):
    matrix = np.zeros((5, 8), np.float32)
    matrix[:, 0] = 1.0
    q = np.zeros(8, np.float32)
    q[0] = -1.0
    idx, sim = orc.fp64_topk(matrix, q, 3, normalized=True)
    assert idx[0].tolist() == [0, 1, 2]
    assert sim[0].tolist() == [-1.0, -1.0, -1.0]
    # normalized=True uses values as given: a non-unit row keeps its length
    matrix[2, 0] = -3.0
    _, s = orc.fp64_topk(matrix, q, 1, normalized=True)
    assert s[0, 0] == 3.0


def test_fp64_matches_rotation_invariance(orc):
    """Cosine scores are invariant under a common orthogonal rotation."""
    rng = np.random.default_rng(5)
    m = rng.normal(size=(300, 16)).astype(np.float32)
    q = rng.normal(size=(16,)).astype(np.float32)
    rot, _ = np.linalg.qr(rng.normal(size=(16, 16)))
    s0 = orc.fp64_scores(m, q, normalized=False)
    s1 = orc.fp64_scores((m.astype(np.float64) @ rot).astype(np.float32),
                         (q.astype(np.float64) @ rot).astype(np.float32), normalized=False)
    np.testing.assert_allclose(s0, s1, atol=1e-6)


# ---------------------------------------------------------------------------
# Normalisation KATs: src/lib.rs:1301-1352, 1144-1182
# ---------------------------------------------------------------------------

def test_normalize_vector_unit_and_padding(orc):
    v = orc.normalize_vector(np.array([3.0, 4.0], np.float32), 8)
    assert v.shape == (8,)
    assert v[:2].tolist() == [np.float32(0.6), np.float32(0.8)]
    assert v[2:].tolist() == [0.0] * 6


def test_normalize_rows_unit(orc):
    vals = np.array([3.0, 4.0, 0.0, 0.0, 6.0, 8.0, 0.0, 0.0], np.float32)
    out = orc.normalize_rows(vals, 2, 4, 8)
    assert out.shape == (2, 8)
    for row in out:
        assert abs(float(np.sqrt((row.astype(np.float64) ** 2).sum())) - 1.0) < 1e-5


def test_pad_rows(orc):
    vals = np.arange(1, 7, dtype=np.float32)
    out = orc.pad_rows(vals, 2, 3, 8)
    assert out[0, :3].tolist() == [1, 2, 3] and out[0, 3:].tolist() == [0] * 5
    assert out[1, :3].tolist() == [4, 5, 6] and out[1, 3:].tolist() == [0] * 5


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_normalize_rejects_non_finite(orc, bad):
    with pytest.raises(orc.OracleError) as e:
        orc.normalize_rows(np.array([1, bad, 0, 0, 0, 0, 0, 0], np.float32), 1, 8, 8)
    assert e.value.kind == "NonFiniteInput"
    with pytest.raises(orc.OracleError):
        orc.normalize_vector(np.array([bad, 0, 0], np.float32))


def test_normalize_rejects_zero_norm(orc):
    with pytest.raises(orc.OracleError) as e:
        orc.normalize_rows(np.zeros(8, np.float32), 1, 8, 8)
    assert e.value.kind == "ZeroNorm"
    with pytest.raises(orc.OracleError) as e:
        orc.normalize_vector(np.zeros(8, np.float32))
    assert e.value.kind == "ZeroNorm"


# ---------------------------------------------------------------------------
# Exact search KATs: src/lib.rs:824-899, 1374-1506
# ---------------------------------------------------------------------------

def test_lib_exact_top_k_matches_expected(orc):
    matrix = np.zeros((3, 8), np.float32)
    matrix[0, 0] = 1
    matrix[1, 1] = 1
    matrix[2, :2] = 1
    m = orc.normalize_rows(matrix, 3, 8)
    q = orc.normalize_vector(np.array([0.8, 0.6, 0, 0, 0, 0, 0, 0], np.float32))
    idx, sim, _ = orc.cpu_search(m, q, 2)
    assert idx.tolist() == [2, 0]
    assert abs(sim[0] - 0.98994946) < 1e-5 and abs(sim[1] - 0.8) < 1e-6


def test_lib_k_larger_than_rows_and_all_negative(orc):
    m = np.zeros((5, 8), np.float32)
    m[:, 0] = 1
    idx, sim, _ = orc.cpu_search(m, np.array([-1] + [0] * 7, np.float32), 100)
    assert len(idx) == 5 and np.all(sim == -1.0)
    bidx, bsim, lens, _ = orc.cpu_search_batch(m, np.tile([-1] + [0] * 7, (4, 1)).astype(
        np.float32), 3)
    assert lens.tolist() == [3] * 4 and np.all(bsim == -1.0)


@pytest.mark.parametrize("rows,dims", [(100, 16), (3000, 48), (20000, 64), (60000, 8)])
def test_strategies_agree_with_serial_dot(orc, rows, dims):
    """Every restated strategy returns the textbook top-k of the same dots."""
    rng = np.random.default_rng(rows + dims)
    m = orc.normalize_rows(rng.uniform(-1, 1, size=(rows, dims)).astype(np.float32), rows, dims)
    q = orc.normalize_vector(rng.uniform(-1, 1, size=dims).astype(np.float32))
    scores = np.array([orc.dot_product(r, q) for r in m], np.float32)
    order = sorted(range(rows), key=lambda i: (-scores[i], i))[:16]
    for threads in (1, 3, 8):
        idx, sim, _ = orc.cpu_search(m, q, 16, threads=threads)
        assert idx.tolist() == order
        assert sim.tolist() == scores[order].tolist()


# ---------------------------------------------------------------------------
# topk.rs KATs (src/topk.rs:216-477)
# ---------------------------------------------------------------------------

def _acc(orc, k, pushes):
    a = orc.TopK(k)
    for i, s in pushes:
        a.push(i, s)
    return a


def test_topk_retains_top_k(orc):
    idx, _ = _acc(orc, 3, [(0, .1), (1, .9), (2, .5), (3, .3), (4, .7)]).into_sorted()
    assert idx.tolist() == [1, 4, 2]


def test_topk_fast_reject(orc):
    idx, _ = _acc(orc, 2, [(0, .8), (1, .9), (2, .1), (3, 0.), (4, -1.)]).into_sorted()
    assert idx.tolist() == [1, 0]


def test_topk_merge(orc):
    left = _acc(orc, 3, [(0, .9), (1, .1), (2, .5)])
    right = _acc(orc, 3, [(3, .8), (4, .2), (5, .7)])
    left.merge(right)
    assert left.into_sorted()[0].tolist() == [0, 3, 5]


def test_topk_ties_lower_index_first(orc):
    assert _acc(orc, 3, [(5, .5), (2, .5), (8, .5)]).into_sorted()[0].tolist() == [2, 5, 8]


def test_topk_equal_score_rejected_when_full(orc):
    assert sorted(_acc(orc, 2, [(0, .5), (1, .5), (2, .5)]).into_sorted()[0].tolist()) == [0, 1]


def test_topk_heap_evicts_lowest_index_among_equal_minima(orc):
    # (score, index) heap order: the minimum of {(0.5,0),(0.5,1)} is index 0
    assert _acc(orc, 2, [(0, .5), (1, .5), (2, .7)]).into_sorted()[0].tolist() == [2, 1]


def _heap_rule(scores, k):
    """Closed form of a bounded heap fed every row in row order (push rejects
    a score <= the minimum of a full heap; eviction takes the lowest score,
    lowest row first): G = rows above the k-th score s, t* = the k-th smallest
    row among the rows >= s, P = rows tied at s up to t*; the heap keeps G and
    P less its e lowest rows, e = |G rows after t*|."""
    n = len(scores)
    k = min(k, n)
    canon = sorted(range(n), key=lambda r: (-scores[r], r))
    if k <= 1 or n == k:
        return canon[:k]
    s = scores[canon[k - 1]]
    g_rows = [r for r in range(n) if scores[r] > s]
    t_rows = [r for r in range(n) if scores[r] == s]
    tstar = sorted(g_rows + t_rows)[k - 1]
    p_rows = [r for r in t_rows if r <= tstar]
    e = sum(1 for r in g_rows if r > tstar)
    return sorted(g_rows + p_rows[e:], key=lambda r: (-scores[r], r))


def test_gpu_exact_search_hand_worked_tie(orc):
    # scores 0.5, 0.5, 0.9 (exact in f32): the third row evicts row 0, the
    # lowest row among the two minima, so the heap answers [2, 1], not [2, 0]
    m = np.array([[0.5, 0.75 ** 0.5], [0.5, 0.75 ** 0.5], [0.9, 0.19 ** 0.5]], np.float32)
    idx, sim = orc.gpu_exact_search(m, np.array([1.0, 0.0], np.float32), 2)
    assert idx.tolist() == [2, 1] and sim.tolist() == [np.float32(0.9), np.float32(0.5)]


@pytest.mark.parametrize("seed", range(6))
def test_gpu_exact_search_matches_closed_form_heap_rule(orc, seed):
    """Rows drawn from a small pool (exact ties everywhere): the oracle's
    GPU-path search equals the closed form for every k, and differs from the
    (score desc, row asc) order somewhere."""
    rng = np.random.default_rng(seed)
    pool = rng.normal(size=(5, 16)).astype(np.float32)
    pool /= np.linalg.norm(pool, axis=1, keepdims=True)
    m = np.ascontiguousarray(pool[rng.integers(0, 5, size=60)])
    q = pool[0] + 0.3 * rng.normal(size=16).astype(np.float32)
    q = (q / np.linalg.norm(q)).astype(np.float32)
    scores = [orc.dot_product(m[r], q) for r in range(60)]
    differs = False
    for k in range(1, 61):
        idx, sim = orc.gpu_exact_search(m, q, k)
        want = _heap_rule(scores, k)
        assert idx.tolist() == want, k
        assert sim.tolist() == [np.float32(scores[r]) for r in want]
        differs |= want != sorted(range(60), key=lambda r: (-scores[r], r))[:k]
    assert differs


def test_topk_fewer_than_k_and_k1(orc):
    assert _acc(orc, 10, [(0, .3), (1, .7)]).into_sorted()[0].tolist() == [1, 0]
    idx, sim = _acc(orc, 1, [(0, .1), (1, .9), (2, .5)]).into_sorted()
    assert idx.tolist() == [1] and sim.tolist() == [np.float32(0.9)]
    left, right = _acc(orc, 1, [(0, .3)]), _acc(orc, 1, [(9, .7)])
    left.merge(right)
    assert left.into_sorted()[0].tolist() == [9]


def test_topk_negative(orc):
    assert _acc(orc, 2, [(0, -.9), (1, -.1), (2, -.5)]).into_sorted()[0].tolist() == [1, 2]


def test_topk_from_scores_kats(orc):
    assert orc.topk_from_scores([.1, .9, .5, .3, .7], 3)[0].tolist() == [1, 4, 2]
    assert orc.topk_from_scores([.1, .9, .5], 2, 100)[0].tolist() == [101, 102]
    assert orc.topk_from_scores([.3, .7], 10)[0].tolist() == [1, 0]
    assert orc.topk_from_scores([], 5)[0].tolist() == []
    assert orc.topk_from_scores([.5, .5, .5], 3)[0].tolist() == [0, 1, 2]


def test_topk_from_scores_streaming_boundaries(orc):
    s = np.arange(200_000, dtype=np.float32) * np.float32(0.001)
    s[150_000], s[175_000] = 999.0, 998.0
    assert orc.topk_from_scores(s, 3)[0].tolist() == [150_000, 175_000, 199_999]
    a = orc.topk_from_scores(np.arange(131_072, dtype=np.float32), 5)[0]
    assert a[:2].tolist() == [131_071, 131_070]
    b = orc.topk_from_scores(np.arange(131_073, dtype=np.float32), 5)[0]
    assert b[:2].tolist() == [131_072, 131_071]
    s = np.arange(150_000, dtype=np.float32) * np.float32(0.01)
    assert orc.topk_from_scores(s, 2, 1_000_000)[0].tolist() == [1_149_999, 1_149_998]
    full = np.sin(np.arange(140_000, dtype=np.float32) * np.float32(0.0137))
    lo = orc.topk_from_scores(full[:131_072], 10)
    hi = orc.topk_from_scores(full, 10)
    assert len(lo[0]) == len(hi[0]) == 10 and hi[1][0] >= lo[1][0] - 1e-6


# ---------------------------------------------------------------------------
# profiles.rs KATs (src/profiles.rs:194-343) and cpu.rs helpers
# ---------------------------------------------------------------------------

def test_profile_cutoffs(orc):
    assert orc.profile_cutoff(10_000, 100, 200) == 1024
    assert orc.profile_cutoff(1_000, 100, 200) == 2048
    assert orc.profile_cutoff(100, 32, 1) == 8192


@pytest.mark.parametrize("shape,args,expected", [
    ((100, 32), (100, 128, 500, 128, 4096, True), "Serial"),
    ((1000, 768), (1000, 768, 10000, 128, 4096, True), "Gemv"),
    ((100000, 768), (3000, 768, 500, 128, 4096, True), "Gemv"),
    ((100000, 48), (3000, 48, 500, 128, 4096, True), "ParallelScores"),
    ((1000, 768), (1000, 768, 10000, 128, 4096, False), "Serial"),
    ((1000, 32), (1000, 32, 10000, 128, 4096, True), "Serial"),
    ((100000, 768), (10000, 768, 500, 256, 4096, True), "ParallelChunked"),
    ((100000, 768), (10000, 768, 500, 64, 4096, True), "ParallelFold"),
])
def test_search_strategy(orc, shape, args, expected):
    cut = orc.profile_cutoff(shape[0], shape[1], 1)
    assert orc.cpu_search_strategy(cut, *args) == expected


@pytest.mark.parametrize("args,expected", [
    ((18_181, 640, 96, True), "Gemm"),
    ((10_000, 32, 128, True), "Gemm"),
    ((5_000, 24, 20, True), "Gemm"),
    ((1_000, 8, 128, True), "QueryParallel"),
    ((1_000_000, 128, 100, True), "GemmChunked"),
    ((10_000_000, 8, 5, True), "MatrixChunked"),
    ((1_000, 8, 10, True), "QueryParallel"),
    ((1_000, 8, 2, False), "ProfiledPerQuery"),
])
def test_batch_strategy(orc, args, expected):
    assert orc.cpu_batch_strategy(*args) == expected


def test_cpu_helpers(orc):
    a = np.arange(1, 11, dtype=np.float32)
    b = a[::-1].copy()
    assert abs(orc.dot_unrolled_8(a, b) - float((a * b).sum())) < 1e-6
    assert orc.dot_unrolled_8([3, 4], [1, 2]) == 11.0
    assert orc.dot_unrolled_8([], []) == 0.0
    assert orc.dot_unrolled_8([1, 2, 3], [4, 5, 6, 7, 8]) == 32.0
    a64 = np.arange(64, dtype=np.float32) * np.float32(0.1)
    assert abs(orc.dot_product(a64, a64[::-1].copy()) - float(
        (a64.astype(np.float64) * a64[::-1]).sum())) < 0.01
    assert orc.chunk_rows(1, 100_000) >= 1 and orc.chunk_rows(10, 8) <= 10
    assert orc.batch_chunk_rows(100_000) >= 64
    assert orc.batch_chunk_rows(32) > orc.batch_chunk_rows(1024)
    # config shapes: 1M x 384 single query picks ParallelFold (chunk_rows 85 < 128)
    assert orc.chunk_rows(1_000_000, 384) == 85


# ---------------------------------------------------------------------------
# approx.rs KATs (src/approx.rs:557-731)
# ---------------------------------------------------------------------------

def test_ivf_sizing(orc):
    assert [orc.isqrt(n) for n in (0, 1, 4, 10_000, 65_536)] == [0, 1, 2, 100, 256]
    assert orc.choose_nlist(100) == 1 and orc.choose_nlist(200) == 1
    assert orc.choose_nlist(10_000) == 100 and orc.choose_nlist(1_000_000) == 512
    assert orc.choose_nlist(300) >= 16
    for nlist in (16, 64, 128, 256, 512):
        for k in (1, 5, 10, 20, 50, 100):
            for rows in (1_000, 10_000, 100_000):
                assert 1 <= orc.choose_nprobe(nlist, k, rows) <= nlist


def test_ivf_build_layout(orc):
    rows, dims = 1000, 32
    m = np.sin(np.arange(rows * dims, dtype=np.float32) * np.float32(0.017)).reshape(rows, dims)
    ivf = orc.Ivf(m)
    perm = ivf.original_indices()
    assert sorted(perm.tolist()) == list(range(rows))
    off = ivf.offsets()
    assert off[0] == 0 and off[-1] == rows and np.all(np.diff(off.astype(np.int64)) >= 0)
    small = orc.Ivf(np.ones((100, 32), np.float32))
    assert not small.provides_speedup
    big = orc.Ivf(np.sin(np.arange(4096 * 64, dtype=np.float32) * np.float32(0.013)).reshape(
        4096, 64))
    assert big.provides_speedup


def test_ivf_recall_floor(orc):
    """tests_python/test_accuracy_vs_numpy.py:29-41 (numpy data, reproducible):
    approx top-25 overlaps exact in >= 15.  The ChaCha8-seeded floor of
    src/lib.rs:961-1021 cannot be regenerated without rand_chacha."""
    rng = np.random.default_rng(456)
    rows, dims, k = 4096, 128, 25
    matrix = rng.normal(size=(rows, dims)).astype(np.float32)
    query = rng.normal(size=(dims,)).astype(np.float32)
    m = orc.normalize_rows(matrix, rows, dims)
    q = orc.normalize_vector(query)
    exact, _, _ = orc.cpu_search(m, q, k)
    ivf = orc.Ivf(m)
    assert ivf.nlist == 64 and orc.choose_nprobe(64, k, rows) == 16
    approx, _ = ivf.search(q, k)
    assert len(set(exact.tolist()) & set(approx.tolist())) >= 15
    assert int(approx[0]) in set(exact.tolist())


def test_ivf_search_gpu_path(orc):
    """The GPU backend's approximate search (src/gpu.rs:448-462).  Probing every
    list scans every row once, so the result is the brute-force top-k of the
    simsimd dot over all rows (ordered by score desc, then row asc); with one
    probe every result comes from the cluster whose centroid scores highest."""
    rng = np.random.default_rng(11)
    rows, dims, k = 3000, 40, 10
    m = orc.normalize_rows(rng.normal(size=(rows, dims)).astype(np.float32), rows, dims)
    q = orc.normalize_vector(rng.normal(size=dims).astype(np.float32))
    ivf = orc.Ivf(m)
    assert ivf.nlist == orc.choose_nlist(rows) == 54
    scores = np.array([orc.dot_product(m[r], q) for r in range(rows)], np.float32)
    order = np.lexsort((np.arange(rows), -scores.astype(np.float64)))[:k]
    idx, sim = ivf.search_gpu(q, k, nprobe=ivf.nlist)
    assert idx.astype(np.int64).tolist() == order.tolist()
    assert np.array_equal(sim, scores[order])
    cscore = ivf.centroids().astype(np.float64) @ q.astype(np.float64)
    best = int(np.argmax(cscore))
    assert np.sort(cscore)[-1] - np.sort(cscore)[-2] > 1e-4  # no near-tie on this seed
    off, perm = ivf.offsets(), ivf.original_indices()
    members = set(perm[off[best]:off[best + 1]].tolist())
    idx1, sim1 = ivf.search_gpu(q, k, nprobe=1)
    assert len(idx1) == min(k, len(members)) and set(idx1.tolist()) <= members
    assert np.all(np.diff(sim1) <= 0)
    # nprobe = 0 is the reference heuristic
    a, _ = ivf.search_gpu(q, k)
    b, _ = ivf.search_gpu(q, k, nprobe=orc.choose_nprobe(ivf.nlist, k, rows))
    assert np.array_equal(a, b)
    # fewer candidates than k: every member returned
    small = orc.Ivf(m, nlist=400)
    off = small.offsets()
    sizes = np.diff(off.astype(np.int64))
    idx2, _ = small.search_gpu(q, 50, nprobe=1)
    assert len(idx2) <= 50 and len(idx2) in set(sizes.tolist()) | {50}


def test_ivf_kmeans_invariants(orc):
    """spherical_kmeans (approx.rs:434-535) fixed points: each centroid is the
    normalised f64 mean of its rows (zero when empty), and the assignment is
    consistent with offsets and the permutation."""
    rng = np.random.default_rng(5)
    rows, dims = 2000, 24
    m = orc.normalize_rows(rng.normal(size=(rows, dims)).astype(np.float32), rows, dims)
    ivf = orc.Ivf(m, nlist=30, iters=3)
    cent, assign = ivf.centroids(), ivf.assignments().astype(np.int64)
    off, perm = ivf.offsets().astype(np.int64), ivf.original_indices().astype(np.int64)
    for c in range(ivf.nlist):
        rows_c = perm[off[c]:off[c + 1]]
        assert np.all(assign[rows_c] == c) and np.all(np.diff(rows_c) > 0)
        if len(rows_c) == 0:
            assert np.all(cent[c] == 0)
        else:
            mean = m[rows_c].astype(np.float64).mean(axis=0)
            np.testing.assert_allclose(cent[c], mean / np.linalg.norm(mean), atol=2e-7)


def test_accuracy_vs_numpy_exact(orc):
    """tests_python/test_accuracy_vs_numpy.py:15-26: exact top-25 equals the
    numpy float32 reference; scores within 1e-5."""
    rng = np.random.default_rng(123)
    rows, dims, k = 4096, 128, 25
    matrix = rng.normal(size=(rows, dims)).astype(np.float32)
    query = rng.normal(size=(dims,)).astype(np.float32)
    mn = matrix / np.maximum(np.linalg.norm(matrix, axis=1, keepdims=True), 1e-12)
    qn = query / max(float(np.linalg.norm(query)), 1e-12)
    scores = mn @ qn
    top = np.argpartition(scores, -k)[-k:]
    np_idx = top[np.argsort(scores[top])[::-1]]
    idx, sim, _ = orc.cpu_search(orc.normalize_rows(matrix, rows, dims),
                                 orc.normalize_vector(query), k)
    assert np.array_equal(idx.astype(np.int64), np_idx)
    assert np.allclose(sim, scores[np_idx], atol=1e-5)


# ---------------------------------------------------------------------------
# Committed fixtures (tests/golden/make_fixtures.py): FP64 brute force from
# numpy beside the oracle's restatement of the reference paths
# ---------------------------------------------------------------------------

def _fixtures():
    import os

    path = os.path.join(os.path.dirname(__file__), "golden", "fixtures.npz")
    with np.load(path) as f:
        return {k: f[k] for k in f.files}


FIXTURE_CASES = ["dims37", "ties", "negative", "k_gt_rows"]


def test_fixture_fp64_matches_the_published_golden_vector():
    fx = _fixtures()
    assert fx["golden50k/fp64_idx"][0].tolist() == synth.GOLDEN_50K_INDICES
    np.testing.assert_allclose(fx["golden50k/fp64_sim"][0], synth.GOLDEN_50K_SCORES,
                               atol=1e-5, rtol=0)
    assert fx["golden50k/fp64_idx"].shape == (5, 5)


@pytest.mark.parametrize("name", FIXTURE_CASES)
def test_oracle_reproduces_fixtures_and_agrees_with_fp64(orc, name):
    """The oracle's reference-path results equal the committed ones bit for
    bit, and agree with the independent FP64 brute force: identical scores to
    1e-5 position by position, identical indices except among rows whose FP64
    scores tie (duplicates keep the reference heap's choice)."""
    fx = _fixtures()
    matrix, queries = fx[f"{name}/matrix"], fx[f"{name}/queries"]
    k = int(fx[f"{name}/k"])
    rows, dims = matrix.shape
    m = orc.normalize_rows(matrix, rows, dims)
    qb = orc.normalize_rows(queries, queries.shape[0], dims)
    f_idx, f_sim = fx[f"{name}/fp64_idx"], fx[f"{name}/fp64_sim"]
    assert f_idx.shape[1] == min(k, rows)
    for i in range(queries.shape[0]):
        qs = orc.normalize_vector(queries[i])
        s_idx, s_sim = orc.gpu_exact_search(m, qs, k)
        b_idx, b_sim = orc.gpu_exact_search(m, qb[i], k)
        assert s_idx.tolist() == fx[f"{name}/single_idx"][i].tolist()
        assert np.array_equal(s_sim.view(np.uint32), fx[f"{name}/single_sim"][i].view(np.uint32))
        assert b_idx.tolist() == fx[f"{name}/batch_idx"][i].tolist()
        assert np.array_equal(b_sim.view(np.uint32), fx[f"{name}/batch_sim"][i].view(np.uint32))
        assert orc.cpu_search(m, qs, k)[0].tolist() == fx[f"{name}/cpu_idx"][i].tolist()
        np.testing.assert_allclose(s_sim, f_sim[i], atol=1e-5, rtol=0)
        for j, (a, b) in enumerate(zip(s_idx.tolist(), f_idx[i].tolist())):
            if a != b:  # only among FP64-tied rows
                full = np.asarray(matrix, np.float64)
                full /= np.linalg.norm(full, axis=1, keepdims=True)
                q64 = queries[i].astype(np.float64) / np.linalg.norm(queries[i].astype(np.float64))
                assert abs(full[a] @ q64 - full[b] @ q64) < 1e-12, (name, i, j)
    if name == "negative":
        assert (f_sim < 0).all()
    if name == "ties":
        assert any(fx["ties/single_idx"][i].tolist() != f_idx[i].tolist() for i in range(6))


def test_fixture_script_reproduces_committed_fixtures(tmp_path):
    """The committed fixtures are what tests/golden/make_fixtures.py writes."""
    import os
    import subprocess
    import sys

    here = os.path.dirname(__file__)
    out = str(tmp_path / "fx.npz")
    subprocess.run([sys.executable, os.path.join(here, "golden", "make_fixtures.py"), out],
                   check=True, capture_output=True)
    with np.load(out) as a, np.load(os.path.join(here, "golden", "fixtures.npz")) as b:
        assert sorted(a.files) == sorted(b.files)
        for key in a.files:
            if key.endswith("fp64_sim"):  # float64 products may differ in the last bit
                np.testing.assert_allclose(a[key], b[key], rtol=0, atol=1e-12)
            else:
                assert np.array_equal(a[key], b[key]), key


# ---------------------------------------------------------------------------
# IVF seeding and assignment pinned by closed forms (src/approx.rs:357-469):
# the GPU build equals the oracle bit for bit (tests/test_gpu_ivf.py), so
# these checks tie both to what the reference's own rules fix.
# ---------------------------------------------------------------------------

MASK64 = (1 << 64) - 1


def _kmeanspp_closed_form(rows, dims, nlist, first, mult=6364136223846793005, inc=1,
                          seed_mul=2654435761, seed_add=None):
    """k-means++ on signed, permuted unit axes: every unchosen row is at
    distance exactly 1 from every centroid (dot 0), a chosen row at 0, so the
    total is the unchosen count and the pick is the first row whose count of
    unchosen rows so far reaches (state / u64::MAX) * total."""
    seed_add = dims if seed_add is None else seed_add
    state = (rows * seed_mul + seed_add) & MASK64
    chosen = [first]
    for _ in range(1, nlist):
        unchosen = [r not in chosen for r in range(rows)]
        total = float(sum(unchosen))
        if total <= 0.0:
            chosen.append(chosen[-1])
            continue
        state = (state * mult + inc) & MASK64
        threshold = (float(state) / float(MASK64)) * total
        cum, pick = 0.0, rows - 1
        for r in range(rows):
            cum += 1.0 if unchosen[r] else 0.0
            if cum >= threshold:
                pick = r
                break
        chosen.append(pick)
    return chosen


def _signed_axes(rows, dims, seed):
    rng = np.random.default_rng(seed)
    m = np.zeros((rows, dims), np.float32)
    axes = rng.permutation(dims)[:rows]
    m[np.arange(rows), axes] = rng.choice(np.array([-1.0, 1.0], np.float32), size=rows)
    return m


@pytest.mark.parametrize("rows,dims,nlist", [(40, 64, 12), (96, 96, 60), (200, 256, 200)])
def test_kmeanspp_seeding_closed_form(orc, rows, dims, nlist):
    m = _signed_axes(rows, dims, rows + nlist)
    # every row's dot with the mean is the same f32 value: the lowest row wins
    assert orc.pick_medoid_row(m) == 0
    got = [int(r) for r in orc.kmeans_pp_rows(m, nlist)]
    assert got == _kmeanspp_closed_form(rows, dims, nlist, 0)
    assert len(set(got)) == nlist  # distances stay 1 until every row is chosen
    # a one-constant change of the generator or the seed picks other rows
    for kw in ({"mult": 6364136223846793005 + 2}, {"inc": 3}, {"seed_mul": 2654435769},
               {"seed_add": 0}):
        assert _kmeanspp_closed_form(rows, dims, nlist, 0, **kw) != got, kw


def test_kmeanspp_duplicates_when_every_distance_is_zero(orc):
    """total <= 0 (approx.rs:382-386): the last centroid is repeated."""
    m = np.zeros((10, 16), np.float32)
    m[:7, 0] = 1.0   # seven copies of e0
    m[7, 1] = 1.0
    m[8, 2] = 1.0
    m[9, 3] = 1.0
    got = [int(r) for r in orc.kmeans_pp_rows(m, 8)]
    # medoid: e0's dot with the mean (0.7) beats the others (0.1); rows 1-6
    # tie with row 0, the lowest wins
    assert got[0] == 0
    assert sorted(got[1:4]) == [7, 8, 9]
    assert got[4:] == [got[3]] * 4  # every row at distance 0: duplicates
    assert got == _kmeanspp_closed_form(10, 16, 8, 0)[:1] + got[1:]


def test_medoid_is_the_argmax_against_the_mean(orc):
    rng = np.random.default_rng(5)
    for rows, dims in ((3000, 40), (5000, 128)):
        m = orc.normalize_rows(rng.normal(size=(rows, dims)).astype(np.float32), rows, dims)
        mean = (m.astype(np.float64).sum(axis=0) / rows).astype(np.float32).astype(np.float64)
        dots = m.astype(np.float64) @ mean
        order = np.argsort(-dots, kind="stable")
        assert dots[order[0]] - dots[order[1]] > 1e-5  # f32 rounding cannot reorder
        assert orc.pick_medoid_row(m) == int(order[0])
        # a copy of the best row at a lower index ties it exactly: the lowest wins
        lo = int(order[0]) // 2
        m2 = m.copy()
        m2[lo] = m[order[0]]
        mean2 = (m2.astype(np.float64).sum(axis=0) / rows).astype(np.float32).astype(np.float64)
        d2 = m2.astype(np.float64) @ mean2
        best = np.flatnonzero(d2 == d2.max())
        assert lo in best
        assert orc.pick_medoid_row(m2) == min(best)
        # the argmin (a sign mistake) is another row
        assert int(np.argmin(dots)) != int(order[0])


def test_assignment_is_the_strict_argmax_over_seeded_centroids(orc):
    """One Lloyd iteration (approx.rs:447-469): each row's assignment is the
    first centroid with the largest dot (strict >, the lowest id on ties)."""
    rng = np.random.default_rng(11)
    rows, dims, nlist = 4000, 32, 50
    c = rng.normal(size=(30, dims)).astype(np.float32)
    raw = c[rng.integers(0, 30, size=rows)] + 0.4 * rng.normal(size=(rows, dims)).astype(np.float32)
    m = orc.normalize_rows(raw, rows, dims)
    seeds = orc.kmeans_pp_rows(m, nlist).astype(np.int64)
    ivf = orc.Ivf(m, nlist=nlist, iters=1)
    assign = ivf.assignments().astype(np.int64)
    s = m.astype(np.float64) @ m[seeds].astype(np.float64).T
    top = np.argmax(s, axis=1)  # the first of equal maxima
    srt = np.sort(s, axis=1)
    clear = srt[:, -1] - srt[:, -2] > 1e-5
    assert clear.mean() > 0.9
    np.testing.assert_array_equal(assign[clear], top[clear])
    # duplicated centroids (rows whose seeds repeat) tie exactly: lowest id
    m3 = np.zeros((300, 16), np.float32)  # exact unit rows: e0, and e1 every third
    m3[:, 0] = 1.0
    m3[::3, 0] = 0.0
    m3[::3, 1] = 1.0
    seeds3 = orc.kmeans_pp_rows(m3, 6).astype(np.int64)
    assert len(set(seeds3.tolist())) < 6  # duplicates were seeded
    a3 = orc.Ivf(m3, nlist=6, iters=1).assignments().astype(np.int64)
    d3 = m3.astype(np.float64) @ m3[seeds3].astype(np.float64).T
    for r in range(300):
        assert a3[r] == int(np.flatnonzero(d3[r] == d3[r].max())[0])
