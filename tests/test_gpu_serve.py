"""Serving loop: continuous batching gives every request the same result it
gets alone (exact requests bit for bit equal to the exact search, IVF requests
to the IVF search), across mixed kinds, k and sizes, split requests, concurrent
submitters, and exact batches replayed from captured HIP graphs."""

import threading
import time

import numpy as np
import pytest

from nndex_amd import synth
from nndex_amd.serve import EXACT, IVF, Server

from gpu_util import BackendIndex, DevBuf, DeviceSearch, gpu_available

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the HIP path has no CPU fallback")


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


@pytest.fixture(scope="module")
def setup():
    rows, dims = 60_000, 128
    base, bank = synth.mixture_workload(rows=rows, dims=dims, n_queries=3000)
    b = BackendIndex(base, approx=True)
    b.set_ivf(nlist=128)
    b.build_ivf()
    rng = np.random.default_rng(5)
    reqs = []
    for _ in range(48):
        nq = int(np.exp(rng.uniform(0.0, np.log(700))))
        q0 = int(rng.integers(0, bank.shape[0] - nq))
        k = int(rng.choice([1, 10, 100]))
        mode = int(rng.integers(0, 2))
        nprobe = int(rng.choice([4, 16])) if mode == IVF else 0
        reqs.append((bank[q0:q0 + nq], k, mode, nprobe))
    return b, rows, dims, reqs


def _direct(b, q, k, mode, nprobe):
    """The same request through the device entry points, one call."""
    nq = q.shape[0]
    ds = DeviceSearch(b.h, nq, max(k, 10))
    dq = DevBuf.from_array(q)
    di, dsim, dl = DevBuf(nq * k * 8), DevBuf(nq * k * 4), DevBuf(nq * 8)
    if mode == EXACT:
        ds.run(dq, nq, k, di, dsim)
        lens = np.full(nq, k, np.uint64)
    else:
        ds.run_ivf(dq, nq, k, nprobe, di, dsim, dl)
        lens = dl.to_array((nq,), np.uint64)
    return di.to_array((nq, k), np.uint64), dsim.to_array((nq, k), np.float32), lens


def _same(a, b):
    ai, asim, al = a[:3]
    bi, bsim, bl = b[:3]
    assert np.array_equal(al, bl)
    for r in range(ai.shape[0]):
        n = int(al[r])
        assert np.array_equal(ai[r, :n], bi[r, :n]), r
        assert np.array_equal(asim[r, :n].view(np.uint32), bsim[r, :n].view(np.uint32)), r


def test_serve_alone_equals_direct(setup):
    b, rows, dims, reqs = setup
    _case("serve_alone", requests=len(reqs))
    srv = Server(b.h, dims, rows, slots=1, max_batch_queries=2048, max_k=100, ivf_nprobe_max=16)
    for q, k, mode, nprobe in reqs[:12]:
        got = srv.wait(srv.submit(q, k, mode, nprobe))
        _same(got, _direct(b, q, k, mode, nprobe))
    srv.close()


def test_serve_mixed_equals_alone(setup):
    b, rows, dims, reqs = setup
    _case("serve_mixed", requests=len(reqs))
    # alone: eager launches; mixed: exact batches replay captured HIP graphs
    alone = Server(b.h, dims, rows, slots=1, max_batch_queries=2048, max_k=100,
                   ivf_nprobe_max=16, graphs=False)
    ref = [alone.wait(alone.submit(q, k, m, p)) for q, k, m, p in reqs]
    alone.close()
    # three executors, batches capped at 256 rows (large requests split), all
    # requests queued at once: they merge, split and overlap
    srv = Server(b.h, dims, rows, slots=3, max_batch_queries=256, max_k=100, ivf_nprobe_max=16,
                 graphs=True)
    tickets = [srv.submit(q, k, m, p) for q, k, m, p in reqs]
    got = [srv.wait(t) for t in tickets]
    for r, g in zip(ref, got):
        _same(g, r)
    st = srv.stats()
    print(st, flush=True)
    assert st["requests"] == len(reqs)
    assert st["queries"] == sum(q.shape[0] for q, _, _, _ in reqs)
    assert st["slots"] == 3 and st["batches"] >= 1
    assert st["graph_batches"] >= 1
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    srv.close()


def test_serve_concurrent_submitters(setup):
    """Requests arrive from four threads while batches run; results equal the
    direct search."""
    b, rows, dims, reqs = setup
    _case("serve_concurrent")
    srv = Server(b.h, dims, rows, slots=2, max_batch_queries=512, max_k=100, ivf_nprobe_max=16)
    out = {}
    errors = []

    def worker(w):
        try:
            for i in range(w, len(reqs), 4):
                q, k, m, p = reqs[i]
                t = srv.submit(q, k, m, p)
                time.sleep(0.001)
                out[i] = srv.wait(t)
        except Exception as e:  # surfaced below
            errors.append(e)

    th = [threading.Thread(target=worker, args=(w,)) for w in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for i in range(0, len(reqs), 5):
        q, k, m, p = reqs[i]
        _same(out[i], _direct(b, q, k, m, p))
    srv.close()


def test_serve_admission_and_errors(setup):
    b, rows, dims, reqs = setup
    _case("serve_admission")
    # a budget below one executor's footprint is refused
    with pytest.raises(Exception, match="budget"):
        Server(b.h, dims, rows, slots=2, max_batch_queries=2048, hbm_budget_bytes=1 << 20)
    # a budget for about one executor admits one, whatever slots asks for
    one = Server(b.h, dims, rows, slots=1, max_batch_queries=1024, max_k=100)
    fp = one.stats()["slot_bytes"]
    one.close()
    srv = Server(b.h, dims, rows, slots=4, max_batch_queries=1024, max_k=100,
                 hbm_budget_bytes=int(fp * 1.5))
    assert srv.stats()["slots"] == 1
    q = reqs[0][0]
    with pytest.raises(Exception, match="max_k"):
        srv.submit(q, 101)
    with pytest.raises(Exception, match="without IVF"):
        srv.submit(q, 10, IVF, 4)
    srv.close()
