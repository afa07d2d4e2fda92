"""The Python surface and the query path around the kernels, on the GPU.

- Query preparation on the device: single queries in normalize_vector's
  sequential order (src/lib.rs:558-575), batches in normalize_rows' order
  (src/lib.rs:461-553), padding only for an index of normalized rows
  (src/lib.rs:383-385, :405-408); the checks' errors (src/lib.rs:445-450,
  :544-546) come back with the results.
- from_file (src/python_bindings.rs:68-86, src/python_io.rs:80-377): .npy and
  stored .npz members memory-mapped, compressed .npz and .parquet (list and
  fixed-size-list columns) read, each uploaded through pinned staging.
- Inputs as float64 arrays and lists (src/python_bindings.rs:311-400), and
  pandas DataFrame output (src/python_bindings.rs:499-553).
- Serving raw rows: checked and normalised on the device, a bad row failing
  its own request only.
Every result is compared with the oracle's restatement of the reference path.
"""

import numpy as np
import pytest

import nndex_amd
from nndex_amd.serve import EXACT, IVF, Server

from gpu_util import gpu_available

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the HIP path has no CPU fallback")


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


@pytest.fixture(scope="module")
def data():
    rng = np.random.default_rng(31)
    raw = rng.normal(size=(20_000, 100)).astype(np.float32)
    qs = rng.normal(size=(9, 100)).astype(np.float32) * 3.0
    return raw, qs


def _same(idx, sim, oidx, osim):
    assert np.asarray(idx).tolist() == np.asarray(oidx).tolist()
    assert np.array_equal(np.asarray(sim, np.float32).view(np.uint32),
                          np.asarray(osim, np.float32).view(np.uint32))


def test_query_preparation_orders_on_the_device(orc, data):
    raw, qs = data
    _case("query_orders")
    ix = nndex_amd.NNdex(raw, backend="gpu", enable_cache=False)
    m = orc.normalize_rows(raw, *raw.shape)
    single = np.stack([orc.normalize_vector(q) for q in qs])
    batch = orc.normalize_rows(qs, *qs.shape)
    # the two orders round differently for some query
    assert not np.array_equal(single.view(np.uint32), batch.view(np.uint32))
    for i in range(len(qs)):
        idx, sim = ix.search(qs[i], k=12)
        oidx, osim, _ = orc.cpu_search(m, single[i], 12)
        _same(idx, sim, oidx, osim)
    bidx, bsim = ix.search(qs, k=12)
    for i in range(len(qs)):
        oidx, osim, _ = orc.cpu_search(m, batch[i], 12)
        _same(bidx[i], bsim[i], oidx, osim)
    # an index of normalized rows takes queries as they are (padded)
    ixn = nndex_amd.NNdex(m[:, :100].copy(), normalized=True, backend="gpu", enable_cache=False)
    idx, sim = ixn.search(qs[0], k=5)
    oidx, osim, _ = orc.cpu_search(m, orc.pad_rows(qs[0], 1, 100, 104)[0], 5)
    _same(idx, sim, oidx, osim)


def test_device_checks_raise_the_reference_errors(data):
    raw, qs = data
    _case("query_errors")
    ix = nndex_amd.NNdex(raw, backend="gpu", enable_cache=True)
    bad = qs[0].copy()
    bad[7] = np.nan
    with pytest.raises(RuntimeError, match="all input values must be finite"):
        ix.search(bad, k=3)
    with pytest.raises(RuntimeError, match="encountered a zero-norm vector"):
        ix.search(np.zeros(100, np.float32), k=3)
    two = qs[:2].copy()
    two[1, 3] = np.inf
    with pytest.raises(RuntimeError, match="all input values must be finite"):
        ix.search(two, k=3)
    # a failed call caches nothing: the same query searched again still fails,
    # and good queries still answer
    with pytest.raises(RuntimeError, match="all input values must be finite"):
        ix.search(bad, k=3)
    assert ix.search(qs[1], k=3)[0].shape == (3,)


def test_from_file_formats_equal_the_in_memory_index(orc, data, tmp_path):
    raw, qs = data
    _case("from_file")
    import pyarrow as pa
    import pyarrow.parquet as pq

    ref = nndex_amd.NNdex(raw, backend="gpu", enable_cache=False)
    want = ref.search(qs, k=10)
    paths = {}
    np.save(tmp_path / "m.npy", raw)
    paths["npy"] = (str(tmp_path / "m.npy"), None)
    np.savez(tmp_path / "m.npz", emb=raw, ids=np.arange(3))
    paths["npz"] = (str(tmp_path / "m.npz"), "emb")
    np.savez_compressed(tmp_path / "mc.npz", emb=raw)
    paths["npz_compressed"] = (str(tmp_path / "mc.npz"), "emb")
    flat = pa.array(raw.reshape(-1))
    pq.write_table(pa.table({"emb": pa.FixedSizeListArray.from_arrays(flat, 100)}),
                   tmp_path / "f.parquet", row_group_size=7000)
    paths["parquet_fixed"] = (str(tmp_path / "f.parquet"), "emb")
    offs = pa.array(np.arange(0, raw.size + 1, 100, dtype=np.int32))
    pq.write_table(pa.table({"emb": pa.ListArray.from_arrays(offs, flat)}),
                   tmp_path / "l.parquet")
    paths["parquet_list"] = (str(tmp_path / "l.parquet"), "emb")
    for name, (path, key) in paths.items():
        ix = nndex_amd.NNdex.from_file(path, key=key, backend="gpu", enable_cache=False)
        assert ix.rows == 20_000 and ix.dims == 100, name
        got = ix.search(qs, k=10)
        assert np.array_equal(got[0], want[0]), name
        assert np.array_equal(got[1].view(np.uint32), want[1].view(np.uint32)), name
    m = orc.normalize_rows(raw, *raw.shape)
    oidx, osim, _ = orc.cpu_search(m, orc.normalize_rows(qs, *qs.shape)[0], 10)
    _same(want[0][0], want[1][0], oidx, osim)
    with pytest.raises(ValueError, match="key 'nope' not found"):
        nndex_amd.NNdex.from_file(paths["npz"][0], key="nope", backend="gpu")


def test_f64_and_list_inputs_equal_f32(data):
    raw, qs = data
    _case("input_types")
    a = nndex_amd.NNdex(raw, backend="gpu", enable_cache=False)
    b = nndex_amd.NNdex(raw.astype(np.float64), backend="gpu", enable_cache=False)
    small = raw[:300]
    c = nndex_amd.NNdex(small.tolist(), backend="gpu", enable_cache=False)
    d = nndex_amd.NNdex(small, backend="gpu", enable_cache=False)
    q = qs[2]
    ia, sa = a.search(q, k=7)
    for got in (b.search(q, k=7), a.search(q.astype(np.float64), k=7), a.search(q.tolist(), k=7)):
        _same(got[0], got[1], ia, sa)
    _same(*c.search(q.tolist(), k=7), *d.search(q, k=7))
    bq = a.search(qs[:4].tolist(), k=7)
    bf = a.search(qs[:4], k=7)
    assert np.array_equal(bq[0], bf[0]) and np.array_equal(bq[1].view(np.uint32),
                                                           bf[1].view(np.uint32))


def test_dataframe_output(data):
    raw, qs = data
    _case("dataframe")
    import pandas as pd

    df = pd.DataFrame({"id": np.arange(raw.shape[0]) * 3, "tag": [f"r{i}" for i in range(raw.shape[0])]})
    ix = nndex_amd.NNdex(raw, backend="gpu", enable_cache=False)
    idx, sim = ix.search(qs[0], k=6)
    out = ix.search(qs[0], k=6, dataframe=df)
    assert list(out.columns) == ["id", "tag", "similarity"]
    assert out["id"].tolist() == (idx.astype(np.int64) * 3).tolist()
    assert out["similarity"].tolist() == [float(s) for s in sim]
    outs = ix.search(qs[:3], k=4, dataframe=df)
    bidx, bsim = ix.search(qs[:3], k=4)
    assert isinstance(outs, list) and len(outs) == 3
    for i, o in enumerate(outs):
        assert o["tag"].tolist() == [f"r{int(r)}" for r in bidx[i]]
        assert o["similarity"].tolist() == [float(s) for s in bsim[i]]
    with pytest.raises(ValueError, match="row count must match"):
        ix.search(qs[0], k=3, dataframe=df.iloc[:5])


def test_serving_raw_rows(orc, data):
    raw, qs = data
    _case("serve_raw")
    ix = nndex_amd.NNdex(raw, approx=True, backend="gpu", enable_cache=False)
    h = ix.backend_handle
    srv = Server(h, 104, raw.shape[0], slots=2, max_batch_queries=64, max_k=50, ivf_nprobe_max=8)
    rng = np.random.default_rng(4)
    good = [rng.normal(size=(int(n), 100)).astype(np.float32) for n in (1, 5, 30, 70)]
    bad = good[2].copy()
    bad[11] = 0.0  # a zero row
    t_good = [srv.submit_raw(g, 10, EXACT) for g in good]
    t_bad = srv.submit_raw(bad, 10, EXACT)
    t_ivf = srv.submit_raw(good[3], 5, IVF, 8)
    m = orc.normalize_rows(raw, *raw.shape)
    exact = nndex_amd.NNdex(raw, backend="gpu", enable_cache=False)  # the same rows, exact
    for g, t in zip(good, t_good):
        gi, gs, gl, _ = srv.wait(t)
        if g.shape[0] > 1:  # the front-end's batch path on the same rows (a 1-row
            # query there takes the single-query path and its other order)
            fi, fs = exact.search(g, k=10)
            assert np.array_equal(gi, fi) and np.array_equal(gs.view(np.uint32),
                                                             fs.view(np.uint32))
        gn = orc.normalize_rows(g, *g.shape)
        oidx, osim = orc.gpu_exact_search(m, gn[0], 10)
        _same(gi[0], gs[0], oidx, osim)
    with pytest.raises(Exception, match="encountered a zero-norm vector"):
        srv.wait(t_bad)
    gi, gs, gl, _ = srv.wait(t_ivf)
    ivf = orc.Ivf(m)
    gn = orc.normalize_rows(good[3], *good[3].shape)
    for r in (0, 69):
        oidx, osim = ivf.search_gpu(gn[r], 5, 8)
        assert int(gl[r]) == len(oidx)
        _same(gi[r, : len(oidx)], gs[r, : len(oidx)], oidx, osim)
    with pytest.raises(Exception, match="dimensionality"):
        srv.submit_raw(np.ones((2, 90), np.float32), 5)
    srv.close()
