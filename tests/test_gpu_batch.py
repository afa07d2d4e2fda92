"""Batched MFMA path: exact results through the certificate + fallback, and
the BF16 scoring + FP32 rerank form's recall and certificate rate."""

import ctypes

import numpy as np
import pytest

import nndex_amd
from nndex_amd import _lib, synth

from gpu_util import BackendIndex, DevBuf, DeviceSearch, check, gpu_available

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the HIP path has no CPU fallback")


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


def batch_device(b: BackendIndex, ds: DeviceSearch, qs: np.ndarray, k: int, precision: int,
                 overfetch: int = 0):
    nq = qs.shape[0]
    dq = DevBuf.from_array(qs)
    didx, dsim, dfl = DevBuf(nq * k * 8), DevBuf(nq * k * 4), DevBuf(nq)
    check(_lib.lib().nndex_gpu_index_search_batch_device(
        b.h, ds.ws, dq.ptr, nq, k, precision, overfetch, didx.ptr, dsim.ptr, dfl.ptr, ds.stream))
    check(_lib.lib().nndex_stream_synchronize(ds.stream))
    return (didx.to_array((nq, k), np.uint64), dsim.to_array((nq, k), np.float32),
            dfl.to_array((nq,), np.uint8))


def exact_device(b: BackendIndex, ds: DeviceSearch, qs: np.ndarray, k: int):
    nq = qs.shape[0]
    dq = DevBuf.from_array(qs)
    didx, dsim = DevBuf(nq * k * 8), DevBuf(nq * k * 4)
    ds.run(dq, nq, k, didx, dsim)
    return didx.to_array((nq, k), np.uint64), dsim.to_array((nq, k), np.float32)


@pytest.mark.parametrize("rows,dims,nq,k", [
    (50_000, 128, 40, 10),
    (20_000, 384, 300, 5),
    (9_000, 64, 17, 1),
    (30_000, 96, 64, 100),
    (3_000, 8, 33, 20),
    (12_000, 768, 24, 10),   # above 512 padded dims: BF16 scoring, certified or rescanned
    (100_000, 128, 40, 10),  # 65,536 rows or more: BF16 scoring for exact batches
    (70_000, 64, 6, 3),      # five queries or more take the MFMA path
])
def test_exact_batch_through_frontend(orc, rows, dims, nq, k):
    """Batches above 8 queries take the certified MFMA path; every result must
    still equal the exact FP32 search bit for bit."""
    _case("exact_batch", rows=rows, dims=dims, nq=nq, k=k)
    rng = np.random.default_rng(rows + nq)
    m = rng.normal(size=(rows, dims)).astype(np.float32)
    qs = rng.normal(size=(nq, dims)).astype(np.float32)
    index = nndex_amd.NNdex(m, normalized=False, backend="gpu", enable_cache=False)
    bidx, bsim = index.search(qs, k=k)
    om = orc.normalize_rows(m, rows, dims)
    oq = orc.normalize_rows(qs, nq, dims)
    for i in range(nq):
        oidx, osim, _ = orc.cpu_search(om, oq[i], k)
        assert bidx[i].tolist() == oidx.tolist(), i
        assert bsim[i].tolist() == osim.tolist(), i


def test_f32_batch_device_certifies(orc):
    _case("f32_certify")
    base, queries = synth.mixture_workload(rows=200_000, dims=384, n_queries=512)
    b = BackendIndex(base)
    ds = DeviceSearch(b.h, 512, 100)
    gi, gs, fl = batch_device(b, ds, queries, 10, 0)
    ei, es = exact_device(b, ds, queries[:8], 10)
    assert fl.mean() <= 0.01
    for i in range(8):
        if not fl[i]:
            assert gi[i].tolist() == ei[i].tolist() and gs[i].tolist() == es[i].tolist()
    # certified results match the exact scan everywhere we check
    sample = [i for i in range(0, 512, 37) if not fl[i]]
    for i in sample:
        oidx, osim, _ = orc.cpu_search(base, queries[i], 10)
        assert gi[i].tolist() == oidx.tolist() and gs[i].tolist() == osim.tolist()


@pytest.mark.parametrize("dims,k", [(384, 10), (128, 100), (1024, 10)])
def test_bf16_rerank_recall_and_certificate(orc, dims, k):
    """BF16 scoring with 4k over-fetch and exact FP32 rerank: a certified
    query's result is proven identical to the exact FP32 search, so recall@k
    is 1.0 and the scores are the reference's bits for every certified query
    (checked against the HIP exact path for all of them and against the
    oracle's restatement of the reference GPU path for a sample)."""
    _case("bf16", dims=dims, k=k)
    rows = 100_000 if dims < 1024 else 40_000
    base, queries = synth.mixture_workload(rows=rows, dims=dims, n_queries=256)
    b = BackendIndex(base)
    ds = DeviceSearch(b.h, 256, 4 * k)
    gi, gs, fl = batch_device(b, ds, queries, k, 1)
    ei, es = exact_device(b, ds, queries, k)
    cert = fl == 0
    print(f"certified {cert.mean():.3f}", flush=True)
    assert cert.mean() >= 0.9
    recall = np.mean([len(set(gi[i].tolist()) & set(ei[i].tolist())) / k for i in range(256)])
    print(f"recall@{k} over all queries {recall:.5f}", flush=True)
    certified = np.nonzero(cert)[0]
    rc = np.mean([len(set(gi[i].tolist()) & set(ei[i].tolist())) / k for i in certified])
    assert rc == 1.0  # recall@k of the certified queries
    for i in certified:
        assert gi[i].tolist() == ei[i].tolist() and gs[i].tolist() == es[i].tolist()
    for i in certified[::16]:
        oidx, osim = orc.gpu_exact_search(base, queries[i], k)
        assert gi[i].tolist() == oidx.tolist()
        assert np.array_equal(gs[i].view(np.uint32), osim.view(np.uint32))


@pytest.mark.parametrize("dims,k,anisotropic", [(384, 10, False), (128, 100, False),
                                               (768, 100, True), (1000, 10, False)])
def test_fp8_rerank_recall(orc, dims, k, anisotropic):
    """FP8-E4M3 scoring (per-row and per-query scales) with max(4k, k+64)
    over-fetch and exact FP32 rerank: every returned similarity is the exact
    FP32 score of its row (bit for bit), results are sorted, certified
    queries equal the exact search, and recall@k against exact is high."""
    _case("fp8", dims=dims, k=k, anisotropic=anisotropic)
    rows = 60_000
    if anisotropic:
        base = synth.anisotropic(rows, dims, seed=3)
        queries = synth.anisotropic(256, dims, seed=4)
    else:
        base, queries = synth.mixture_workload(rows=rows, dims=dims, n_queries=256)
    pdims = (dims + 7) // 8 * 8
    if pdims != dims:
        base = np.pad(base, ((0, 0), (0, pdims - dims)))
        queries = np.pad(queries, ((0, 0), (0, pdims - dims)))
    b = BackendIndex(base)
    ds = DeviceSearch(b.h, 256, max(4 * k, k + 64))
    gi, gs, fl = batch_device(b, ds, queries, k, 2)
    ei, es = exact_device(b, ds, queries, k)
    cert = fl == 0
    recall = np.mean([len(set(gi[i].tolist()) & set(ei[i].tolist())) / k for i in range(256)])
    print(f"certified {cert.mean():.3f} recall@{k} {recall:.4f}", flush=True)
    assert recall >= 0.97
    for i in np.nonzero(cert)[0]:
        assert gi[i].tolist() == ei[i].tolist() and gs[i].tolist() == es[i].tolist()
    for i in range(0, 256, 17):
        assert np.all(np.diff(gs[i].astype(np.float64)) <= 0)
        for j in range(0, k, max(1, k // 7)):
            row = int(gi[i, j])
            assert gs[i, j] == np.float32(orc.dot_product(base[row], queries[i])), (i, j)


def test_batch_position_independence_mfma():
    _case("mfma_position")
    rng = np.random.default_rng(9)
    base = rng.normal(size=(60_000, 128)).astype(np.float32)
    base /= np.linalg.norm(base, axis=1, keepdims=True)
    qs = rng.normal(size=(200, 128)).astype(np.float32)
    qs /= np.linalg.norm(qs, axis=1, keepdims=True)
    b = BackendIndex(base)
    ds = DeviceSearch(b.h, 200, 10)
    i1, s1 = exact_device(b, ds, qs, 10)
    perm = rng.permutation(200)
    i2, s2 = exact_device(b, ds, qs[perm], 10)
    assert np.array_equal(i1[perm], i2) and np.array_equal(s1[perm], s2)
    i3, s3 = exact_device(b, ds, qs[:4], 10)  # fused scan path for the first four
    assert np.array_equal(i1[:4], i3) and np.array_equal(s1[:4], s3)
    i3, s3 = exact_device(b, ds, qs[:5], 10)  # a five-query MFMA batch
    assert np.array_equal(i1[:5], i3) and np.array_equal(s1[:5], s3)


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_batch_duplicate_rows_ties_at_threshold(orc, precision):
    """64 distinct rows repeated: every sample threshold sits on a score that
    hundreds of rows share, so candidates tie at the threshold (decided by the
    row part of the key), the per-wave candidate queues fill and flush many
    times per row tile and the per-query lists overflow.  Certified results
    equal the exact search, every returned similarity is exact, and the exact
    search (certificate + scan fallback, device and frontend) equals the
    oracle's restatement of the reference GPU path bit for bit."""
    _case("duplicates", precision=precision)
    rng = np.random.default_rng(77)
    distinct = rng.normal(size=(64, 128)).astype(np.float32)
    distinct /= np.linalg.norm(distinct, axis=1, keepdims=True)
    base = np.ascontiguousarray(np.tile(distinct, (625, 1)))  # 40,000 rows
    qs = rng.normal(size=(48, 128)).astype(np.float32)
    qs /= np.linalg.norm(qs, axis=1, keepdims=True)
    b = BackendIndex(base)
    ds = DeviceSearch(b.h, 48, 74)
    gi, gs, fl = batch_device(b, ds, qs, 10, precision)
    ei, es = exact_device(b, ds, qs, 10)
    for i in range(0, 48, 5):
        oidx, osim = orc.gpu_exact_search(base, qs[i], 10)
        assert ei[i].tolist() == oidx.tolist() and es[i].tolist() == osim.tolist(), i
    for i in range(48):
        if not fl[i]:
            assert gi[i].tolist() == ei[i].tolist() and gs[i].tolist() == es[i].tolist(), i
        for j in range(10):
            assert gs[i, j] == np.float32(orc.dot_product(base[int(gi[i, j])], qs[i])), (i, j)
    # through the frontend (uncertified queries are rescanned exactly)
    index = nndex_amd.NNdex(base, normalized=False, backend="gpu", enable_cache=False)
    bidx, bsim = index.search(qs, k=10)
    om = orc.normalize_rows(base, base.shape[0], 128)
    oq = orc.normalize_rows(qs, 48, 128)
    for i in range(48):
        oidx, osim = orc.gpu_exact_search(om, oq[i], 10)
        assert bidx[i].tolist() == oidx.tolist() and bsim[i].tolist() == osim.tolist(), i


@pytest.mark.parametrize("dims", [128, 384])
def test_batch_tile_variants_agree_and_autotune(dims):
    """The batch autotuner's tile-shape variants (rows per wavefront, queries
    per chunk) give identical results and flags for every precision; the
    autotuner reports its timings and a choice per precision."""
    _case("batch_variants", dims=dims)
    base, queries = synth.mixture_workload(rows=60_000, dims=dims, n_queries=300)
    b = BackendIndex(base)
    ds = DeviceSearch(b.h, 300, 64)
    L = _lib.lib()
    for prec in (0, 1, 2):
        out = []
        for v in (0, 1):
            check(L.nndex_gpu_index_set_batch_tiles(b.h, prec, v))
            out.append(batch_device(b, ds, queries, 10, prec))
        for x, y in zip(out[0], out[1]):
            assert np.array_equal(x, y), prec
    buf = ctypes.create_string_buffer(4096)
    check(L.nndex_gpu_index_autotune_batch(b.h, 1, buf, 4096))
    report = buf.value.decode()
    print(report, flush=True)
    assert report.count("chosen") == 3
    ei, es = exact_device(b, ds, queries, 10)
    gi, gs, fl = batch_device(b, ds, queries, 10, 0)
    for i in range(300):
        if not fl[i]:
            assert gi[i].tolist() == ei[i].tolist() and gs[i].tolist() == es[i].tolist()


@pytest.mark.parametrize("rows,dims", [(150_000, 384), (100_000, 768)])
def test_bf16_scan_many_tiles_per_block(rows, dims):
    """More row tiles than blocks (one block per CU loops over tiles, the
    last round split into query-chunk pieces): the BF16 scans with the chunk
    barrier inside the MFMA stream, at either tile shape, return for every
    certified query exactly the fused FP32 scan's result."""
    _case("bf16_many_tiles", rows=rows, dims=dims)
    base, queries = synth.mixture_workload(rows=rows, dims=dims, n_queries=200)
    b = BackendIndex(base)
    ds = DeviceSearch(b.h, 200, 64)
    L = _lib.lib()
    ei, es = exact_device(b, ds, queries, 10)
    for v in (0, 1):
        check(L.nndex_gpu_index_set_batch_tiles(b.h, 1, v))
        gi, gs, fl = batch_device(b, ds, queries, 10, 1)
        ok = fl == 0
        assert ok.mean() > 0.9, (v, ok.mean())
        for i in np.flatnonzero(ok):
            assert gi[i].tolist() == ei[i].tolist(), (v, i)
            assert gs[i].tolist() == es[i].tolist(), (v, i)
    check(L.nndex_gpu_index_set_batch_tiles(b.h, 1, 0))


@pytest.mark.parametrize("dims", [600, 768, 1000])
def test_wide_rows_tile_variants(dims):
    """Above 512 dims (BF16 and FP8 scoring only) with either tile variant:
    every certified BF16 query equals the exact FP32 search bit for bit, and
    FP8 returns exact FP32 scores of the rows it finds, with recall@10 near 1
    on the mixture."""
    _case("wide_rows", dims=dims)
    base, queries = synth.mixture_workload(rows=40_000, dims=dims, n_queries=120)
    b = BackendIndex(base)
    ds = DeviceSearch(b.h, 120, 64)
    L = _lib.lib()
    ei, es = exact_device(b, ds, queries, 10)
    for prec in (1, 2):
        for v in (0, 1):
            check(L.nndex_gpu_index_set_batch_tiles(b.h, prec, v))
            gi, gs, fl = batch_device(b, ds, queries, 10, prec)
            if prec == 1:
                ok = fl == 0
                assert ok.mean() > 0.9, (prec, v, ok.mean())
                for i in np.flatnonzero(ok):
                    assert gi[i].tolist() == ei[i].tolist(), (prec, v, i)
                    assert gs[i].tolist() == es[i].tolist(), (prec, v, i)
            else:
                hits = sum(len(set(gi[i].tolist()) & set(ei[i].tolist())) for i in range(120))
                assert hits / 1200 > 0.99, (v, hits)
                exact = {(i, int(r)): s for i in range(120) for r, s in zip(ei[i], es[i])}
                for i in range(120):
                    for r, s in zip(gi[i], gs[i]):
                        if (i, int(r)) in exact:
                            assert s == exact[(i, int(r))], (v, i, int(r))
    check(L.nndex_gpu_index_set_batch_tiles(b.h, 1, 0))
    check(L.nndex_gpu_index_set_batch_tiles(b.h, 2, 0))


def _fp8_values():
    """The 256 OCP E4M3 code values (NaN codes 0x7F / 0xFF as nan)."""
    c = np.arange(256)
    e, m = (c >> 3) & 15, c & 7
    v = np.where(e == 0, m * 2.0 ** -9, (1 + m / 8.0) * 2.0 ** (e - 7.0))
    v = np.where((c & 0x7F) == 0x7F, np.nan, v)
    return np.where(c & 0x80, -v, v)


@pytest.mark.parametrize("elem", [0, 1, 2])
def test_mfma_accumulation_within_certificate_bounds(elem):
    """The certificate bounds |approximate - exact| of every MFMA score by
    gamma_K sum|a b| for the FP32 and BF16 forms (an FP32 sum) and by
    kFp8MfmaRelErr = 2^-8 of sum|a b| per 128-product instruction, plus FP32
    accumulation over the chain, for FP8 (whose MFMA does not sum within
    FP32's bound).  Random operands of narrow and wide exponent ranges, one
    instruction and chains of 12, against exact float64 sums."""
    L = _lib.lib()
    rng = np.random.default_rng(100 + elem)
    u = 2.0 ** -24
    f8 = _fp8_values()
    worst = 0.0
    for trial in range(24):
        calls = 1 if trial % 2 == 0 else 12
        wide = trial % 4 >= 2
        if elem == 0:
            a = (rng.uniform(-1, 1, (calls, 64)) * (2.0 ** rng.integers(-20, 20, (calls, 64)) if wide else 1)).astype(np.float32)
            b = (rng.uniform(-1, 1, (calls, 64)) * (2.0 ** rng.integers(-20, 20, (calls, 64)) if wide else 1)).astype(np.float32)
            av, bv = a.astype(np.float64)[..., None], b.astype(np.float64)[..., None]
            per_call = 4
        elif elem == 1:
            fa = (rng.uniform(-1, 1, (calls, 64, 8)) * (2.0 ** rng.integers(-20, 20, (calls, 64, 8)) if wide else 1)).astype(np.float32)
            fb = (rng.uniform(-1, 1, (calls, 64, 8)) * (2.0 ** rng.integers(-20, 20, (calls, 64, 8)) if wide else 1)).astype(np.float32)
            a = (fa.view(np.uint32) >> 16).astype(np.uint16)
            b = (fb.view(np.uint32) >> 16).astype(np.uint16)
            av = (a.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
            bv = (b.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
            per_call = 32
        else:
            lo = 0 if wide else 0x30
            hi = 0x7E if wide else 0x60
            a = (rng.integers(lo, hi, (calls, 64, 32)) | (rng.integers(0, 2, (calls, 64, 32)) << 7)).astype(np.uint8)
            b = (rng.integers(lo, hi, (calls, 64, 32)) | (rng.integers(0, 2, (calls, 64, 32)) << 7)).astype(np.uint8)
            av, bv = f8[a], f8[b]
            per_call = 128
        out = np.zeros(256, np.float32)
        check(L.nndex_debug_mfma_tile(elem, a.ctypes.data, b.ctypes.data, calls, out.ctypes.data))
        got = out.reshape(16, 16).astype(np.float64)
        # lane l holds row / column l % 16 and K block l / 16
        A = av.reshape(calls, 4, 16, -1).transpose(2, 0, 1, 3).reshape(16, -1)
        B = bv.reshape(calls, 4, 16, -1).transpose(2, 0, 1, 3).reshape(16, -1)
        exact = A @ B.T
        absum = np.abs(A) @ np.abs(B).T
        K = per_call * calls
        if elem == 2:
            bound = (2.0 ** -8 + (calls + 1) * u / (1 - (calls + 1) * u)) * absum
        else:
            bound = K * u / (1 - K * u) * absum
        err = np.abs(got - exact)
        assert np.all(err <= bound), (trial, float((err / np.maximum(absum, 1e-300)).max()))
        worst = max(worst, float((err / np.maximum(absum, 1e-300)).max()))
    print(f"form {elem}: largest |err| / sum|ab| {worst:.3e}", flush=True)
