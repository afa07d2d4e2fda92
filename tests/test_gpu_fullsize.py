"""The benchmark configurations at their own sizes, on the GPU.

- Exact FP32 1,000,000 x 384 Gaussian mixture, k=10: the single-query fused
  scan (one query per launch, the headline kernel) on 8 bank queries, and the
  front-end at B=4096, against the oracle's restatement of the reference CPU
  path (bit for bit) and FP64.
- IVF: the 10,000,000 x 128 mixture of 4,096 centres (spread 0.3); its first
  200,000 rows with the reference's own settings (nlist = choose_nlist = 447,
  12 iterations, nprobe = choose_nprobe) against orc.Ivf bit for bit; the full
  base with nlist 4,096 and 20 iterations checked through size-independent
  properties (a permutation, offsets, centroids the normalised means of their
  rows, recall@10 against the exact search at nprobe 8 / 32 / 128).
- Precision form: 5,000,000 x 768 anisotropic base, 1,024 queries, k=100,
  over-fetch 400: the exact search against the oracle and FP64 on a sample,
  BF16 and FP8 scoring + exact rerank at recall@100 = 1 against it, certified
  BF16 queries identical to it, every returned score the exact FP32 one.
- Serving: 65,536 x 1,024 (exact batches score on BF16, the k-means
  assignment runs on BF16), mixed exact / IVF requests served alone and all at
  once, equal to each other and to the oracle.
"""

import numpy as np
import pytest

import nndex_amd
from nndex_amd import synth
from nndex_amd.serve import EXACT, IVF, Server

from gpu_util import BackendIndex, DevBuf, DeviceSearch, assert_topk_parity, gpu_available

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the HIP path has no CPU fallback")


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


@pytest.fixture(scope="module")
def c2():
    return synth.mixture_workload(n_queries=4096)


def test_c2_single_query_kernel(orc, c2):
    base, qs = c2
    _case("c2_single", rows=base.shape[0], dims=base.shape[1])
    b = BackendIndex(base)
    ds = DeviceSearch(b.h, 1, 10)
    di, dsim = DevBuf(10 * 8), DevBuf(10 * 4)
    for i in range(8):
        dq = DevBuf.from_array(qs[i:i + 1])
        ds.run(dq, 1, 10, di, dsim)
        gi, gs = di.to_array((10,), np.uint64), dsim.to_array((10,), np.float32)
        oidx, osim, strategy = orc.cpu_search(base, qs[i], 10)
        assert strategy == "ParallelFold"
        assert gi.tolist() == oidx.tolist()
        assert np.array_equal(gs.view(np.uint32), osim.view(np.uint32))
        if i < 3:
            full = orc.fp64_scores(base, qs[i], normalized=True)
            fidx, _ = orc.fp64_topk(base, qs[i], 10, normalized=True)
            assert_topk_parity(gi, gs, fidx[0], full)
    # the benchmark's launch shape (20 tiles per wave: 196 blocks, whose
    # final merge finds its bound by ballots over the block maxima; the
    # default shape above has more than 512 blocks and skips that step)
    from nndex_amd import _lib
    _lib.check(_lib.lib().nndex_gpu_index_set_launch(b.h, 20, 1))
    ds20 = DeviceSearch(b.h, 1, 10)
    for i in range(8):
        dq = DevBuf.from_array(qs[i:i + 1])
        ds20.run(dq, 1, 10, di, dsim)
        gi, gs = di.to_array((10,), np.uint64), dsim.to_array((10,), np.float32)
        oidx, osim, _ = orc.cpu_search(base, qs[i], 10)
        assert gi.tolist() == oidx.tolist(), i
        assert np.array_equal(gs.view(np.uint32), osim.view(np.uint32)), i


def test_c2_batch_4096_through_frontend(orc, c2):
    base, qs = c2
    _case("c2_b4096")
    index = nndex_amd.NNdex(base, normalized=True, backend="gpu", enable_cache=False)
    bidx, bsim = index.search(qs, k=10)
    assert bidx.shape == (4096, 10)
    # normalized=True: queries are checked and padded, not normalised
    # (src/lib.rs:383-385), so the oracle gets them as they are
    qn = qs
    for i in range(0, 4096, 256):  # 16 queries against the oracle
        oidx, osim, _ = orc.cpu_search(base, qn[i], 10)
        assert bidx[i].tolist() == oidx.tolist(), i
        assert np.array_equal(bsim[i].view(np.uint32), osim.view(np.uint32)), i
    assert np.all(np.diff(bsim, axis=1) <= 0)
    assert all(len(set(r.tolist())) == 10 for r in bidx)
    for i in range(1, 4096, 97):  # reported scores are the rows' reference dots
        direct = [orc.dot_product(base[int(r)], qn[i]) for r in bidx[i]]
        assert direct == bsim[i].tolist(), i


def _c4(rows):
    centres = synth.mixture_centres(128, 4096, seed=7)
    return centres, synth.gaussian_mixture(rows, 128, centres, 0.3, 42)


def test_c4_first_200k_rows_equal_the_oracle(orc):
    """The IVF base's first 200,000 rows (one generator block, cut) with the
    reference's own settings: build and search identical to the oracle."""
    _case("c4_200k")
    centres, block = _c4(1 << 18)
    m = np.ascontiguousarray(block[:200_000])
    qs = synth.gaussian_mixture(64, 128, centres, 0.3, 43)
    b = BackendIndex(m, approx=True)
    info, cent, off, orig, assign = b.ivf_export()
    ref = orc.Ivf(m)
    assert int(info.nlist) == ref.nlist == 447 and int(info.iters) == 12
    np.testing.assert_array_equal(assign, ref.assignments())
    np.testing.assert_array_equal(off, ref.offsets())
    np.testing.assert_array_equal(orig, ref.original_indices())
    assert np.array_equal(cent.view(np.uint32), ref.centroids().view(np.uint32))
    bidx, bsim, lens = b.search_batch(qs, 10, approx=True)
    for i, q in enumerate(qs):
        ridx, rsim = ref.search_gpu(q, 10)
        assert int(lens[i]) == len(ridx)
        np.testing.assert_array_equal(bidx[i, : lens[i]], ridx)
        assert np.array_equal(bsim[i, : lens[i]].view(np.uint32), rsim.view(np.uint32))


def test_c4_full_10m_build_properties_and_recall():
    rows, nlist = 10_000_000, 4096
    _case("c4_10m", rows=rows, nlist=nlist)
    centres, m = _c4(rows)
    qs = synth.gaussian_mixture(256, 128, centres, 0.3, 43)
    b = BackendIndex(m, approx=True)
    b.set_ivf(nlist=nlist, iters=20)
    info, cent, off, orig, assign = b.ivf_export()
    print(f"build {info.build_ms:.0f} ms, rechecked {info.rechecked}", flush=True)
    assert int(info.nlist) == nlist and int(info.iters) == 20
    o = orig.astype(np.int64)
    assert np.array_equal(np.sort(o), np.arange(rows))
    off64 = off.astype(np.int64)
    assert off64[0] == 0 and off64[-1] == rows and np.all(np.diff(off64) >= 0)
    assert np.array_equal(np.repeat(np.arange(nlist), np.diff(off64)), assign[o].astype(np.int64))
    # rows inside a list stay in row order (the stable reorder of approx.rs:64-89)
    starts = off64[:-1][np.diff(off64) > 1]
    for s0 in starts[::97]:
        c = int(assign[o[s0]])
        seg = o[off64[c]:off64[c + 1]]
        assert np.all(np.diff(seg) > 0)
    for c in range(0, nlist, 211):  # the last update: normalised f64 means
        members = o[off64[c]:off64[c + 1]]
        if len(members) == 0:
            assert np.all(cent[c] == 0)
            continue
        mean = m[members].astype(np.float64).mean(axis=0)
        np.testing.assert_allclose(cent[c], mean / np.linalg.norm(mean), atol=2e-7)
    del m
    eidx = b.search_batch(qs, 10)[0]
    for nprobe, floor in ((8, 0.98), (32, 0.99), (128, 0.995)):
        b.set_ivf(nprobe=nprobe)
        aidx = b.search_batch(qs, 10, approx=True)[0]
        rec = float(np.mean([len(set(aidx[i].tolist()) & set(eidx[i].tolist())) / 10
                             for i in range(len(qs))]))
        print(f"nprobe {nprobe}: recall@10 {rec:.4f}", flush=True)
        assert rec >= floor


def test_serving_at_1024_dims_alone_and_mixed(orc):
    """65,536 x 1,024: exact batches score on BF16 under the certificate and
    the k-means assignment runs on BF16 with the exact re-check; requests of
    both kinds give the same rows served alone and all at once, equal to the
    oracle's restatements of the reference GPU path."""
    rows, dims = 65_536, 1024
    _case("serve_1024", rows=rows, dims=dims)
    base, bank = synth.mixture_workload(rows=rows, dims=dims, n_queries=2000)
    b = BackendIndex(base, approx=True)
    b.set_ivf(nlist=64, iters=4)
    b.build_ivf()
    rng = np.random.default_rng(9)
    reqs = []
    for _ in range(24):
        nq = int(np.exp(rng.uniform(0.0, np.log(600))))
        q0 = int(rng.integers(0, bank.shape[0] - nq))
        k = int(rng.choice([1, 10, 100]))
        mode = int(rng.integers(0, 2))
        reqs.append((bank[q0:q0 + nq], k, mode, 32 if mode == IVF else 0))
    alone = Server(b.h, dims, rows, slots=1, max_batch_queries=2048, max_k=100,
                   ivf_nprobe_max=32)
    ref = [alone.wait(alone.submit(q, k, md, p)) for q, k, md, p in reqs]
    alone.close()
    srv = Server(b.h, dims, rows, slots=3, max_batch_queries=256, max_k=100, ivf_nprobe_max=32)
    got = [srv.wait(t) for t in [srv.submit(q, k, md, p) for q, k, md, p in reqs]]
    st = srv.stats()
    srv.close()
    print(st, flush=True)
    assert st["batches"] >= 2
    for (ri, rs, rl, _), (gi, gs, gl, _) in zip(ref, got):
        assert np.array_equal(rl, gl)
        for r in range(ri.shape[0]):
            n = int(rl[r])
            assert np.array_equal(ri[r, :n], gi[r, :n])
            assert np.array_equal(rs[r, :n].view(np.uint32), gs[r, :n].view(np.uint32))
    ivf = orc.Ivf(base, nlist=64, iters=4)
    checked = 0
    for (q, k, md, p), (gi, gs, gl, _) in zip(reqs, got):
        for r in (0, q.shape[0] - 1):
            if md == EXACT:
                oi, osim = orc.gpu_exact_search(base, q[r], k)
            else:
                oi, osim = ivf.search_gpu(q[r], k, p)
            assert int(gl[r]) == len(oi)
            np.testing.assert_array_equal(gi[r, : len(oi)], oi)
            assert np.array_equal(gs[r, : len(oi)].view(np.uint32), osim.view(np.uint32))
            checked += 1
    assert checked >= 24


def test_c3_precision_form_recall(orc):
    """5M x 768 anisotropic, 1,024 queries, k = 100, 4k over-fetch: BF16 and
    FP8 scoring + exact FP32 rerank return the exact search's results
    (recall@100 = 1 on this workload), certified BF16 queries bit for bit;
    the exact search equals the oracle on a sample."""
    _case("c3", rows=5_000_000, dims=768)
    base = synth.anisotropic(5_000_000, 768, seed=42)
    qs = synth.anisotropic(1024, 768, seed=43)
    k, nq = 100, 1024
    b = BackendIndex(base)
    ds = DeviceSearch(b.h, nq, 448)
    dq = DevBuf.from_array(qs)
    ei, es = DevBuf(nq * k * 8), DevBuf(nq * k * 4)
    ds.run(dq, nq, k, ei, es)
    exact_i, exact_s = ei.to_array((nq, k), np.uint64), es.to_array((nq, k), np.float32)
    for i in (0, 511):
        oidx, osim, _ = orc.cpu_search(base, qs[i], k)
        assert exact_i[i].tolist() == oidx.tolist()
        assert np.array_equal(exact_s[i].view(np.uint32), osim.view(np.uint32))
    from nndex_amd import _lib
    L = _lib.lib()
    for prec, name in ((1, "bf16"), (2, "fp8")):
        di, dsim, dfl = DevBuf(nq * k * 8), DevBuf(nq * k * 4), DevBuf(nq)
        from gpu_util import check
        check(L.nndex_gpu_index_search_batch_device(b.h, ds.ws, dq.ptr, nq, k, prec, 4 * k,
                                                    di.ptr, dsim.ptr, dfl.ptr, ds.stream))
        check(L.nndex_stream_synchronize(ds.stream))
        gi, gs = di.to_array((nq, k), np.uint64), dsim.to_array((nq, k), np.float32)
        fl = dfl.to_array((nq,), np.uint8)
        rq = np.array([len(set(gi[i].tolist()) & set(exact_i[i].tolist())) / k for i in range(nq)])
        print(f"{name}: recall@{k} {rq.mean():.5f}, certified {1 - fl.mean():.3f}", flush=True)
        assert rq.min() == 1.0, (name, int((rq < 1).sum()), float(rq.min()))
        for i in np.nonzero(fl == 0)[0]:
            assert gi[i].tolist() == exact_i[i].tolist()
            assert np.array_equal(gs[i].view(np.uint32), exact_s[i].view(np.uint32))
        # every returned score is its row's exact FP32 score: it equals the
        # exact search's score of the same row
        for i in range(0, nq, 97):
            pos = {int(r): j for j, r in enumerate(exact_i[i].tolist())}
            for j in range(k):
                assert gs[i, j] == exact_s[i, pos[int(gi[i, j])]]
        if name == "bf16":
            assert (fl == 0).mean() >= 0.9
    del base
