"""Writes tests/golden/fixtures.npz: small seeded search cases with their
expected results, for the oracle and GPU parity tests.

Expected values come from two independent sources computed here:
  * FP64 brute force in numpy (rows and queries normalised in float64, scores
    in float64, order = score descending then row ascending);
  * the oracle's restatement of the reference paths on the same inputs: the
    GPU backend's exact search (src/gpu.rs:575-589, one bounded heap over the
    rows in order) and the CPU path (src/cpu.rs), both after the reference's
    f32 normalisation (src/lib.rs:461-599: normalize_rows for the matrix and
    batch queries, normalize_vector for a single query).

Cases (each at most 4096 x 128): dims not a multiple of 8, duplicated rows
(ties across the k-th place), all-negative scores, k above the row count, and
the expected FP64 top-5 of the 50,000 x 128 case of
notebooks/python_bindings_demo.ipynb:33-37 (its inputs are regenerated from
numpy's default_rng(42) at test time, so only the results are stored).

Run from the repository root: python tests/golden/make_fixtures.py [out.npz]
"""

from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)

import oracle as orc  # noqa: E402
from nndex_amd import synth  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures.npz")


def fp64_topk(matrix: np.ndarray, queries: np.ndarray, k: int):
    m = matrix.astype(np.float64)
    m /= np.linalg.norm(m, axis=1, keepdims=True)
    q = queries.astype(np.float64)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    s = q @ m.T
    kk = min(k, m.shape[0])
    idx = np.empty((q.shape[0], kk), np.int64)
    sim = np.empty((q.shape[0], kk), np.float64)
    rows = np.arange(m.shape[0])
    for i in range(q.shape[0]):
        order = np.lexsort((rows, -s[i]))[:kk]
        idx[i], sim[i] = order, s[i, order]
    return idx, sim


def cases():
    rng = np.random.default_rng(2024)
    out = {}
    # dims not a multiple of 8 (padded to 40)
    out["dims37"] = (rng.normal(size=(1500, 37)).astype(np.float32),
                     rng.normal(size=(6, 37)).astype(np.float32), 10)
    # 64 distinct rows repeated: many rows tie across the k-th place
    pool = rng.normal(size=(64, 24)).astype(np.float32)
    tied = pool[rng.integers(0, 64, size=1200)]
    tq = pool[rng.integers(0, 64, size=6)] + 0.3 * rng.normal(size=(6, 24)).astype(np.float32)
    out["ties"] = (tied, tq.astype(np.float32), 20)
    # every score negative
    out["negative"] = (-np.abs(rng.normal(size=(800, 16))).astype(np.float32),
                       np.abs(rng.normal(size=(6, 16))).astype(np.float32), 7)
    # k above the row count (capped to rows)
    out["k_gt_rows"] = (rng.normal(size=(40, 8)).astype(np.float32),
                        rng.normal(size=(6, 8)).astype(np.float32), 100)
    return out


def main(out: str = OUT) -> None:
    arrays = {}
    for name, (matrix, queries, k) in cases().items():
        rows, dims = matrix.shape
        nq = queries.shape[0]
        m = orc.normalize_rows(matrix, rows, dims)
        qb = orc.normalize_rows(queries, nq, dims)
        kk = min(k, rows)
        f_idx, f_sim = fp64_topk(matrix, queries, k)
        single_idx = np.zeros((nq, kk), np.uint64)
        single_sim = np.zeros((nq, kk), np.float32)
        batch_idx = np.zeros((nq, kk), np.uint64)
        batch_sim = np.zeros((nq, kk), np.float32)
        cpu_idx = np.zeros((nq, kk), np.uint64)
        for i in range(nq):
            qs = orc.normalize_vector(queries[i])
            single_idx[i], single_sim[i] = orc.gpu_exact_search(m, qs, k)
            batch_idx[i], batch_sim[i] = orc.gpu_exact_search(m, qb[i], k)
            cpu_idx[i] = orc.cpu_search(m, qs, k)[0]
        arrays.update({
            f"{name}/matrix": matrix, f"{name}/queries": queries,
            f"{name}/k": np.int64(k),
            f"{name}/fp64_idx": f_idx, f"{name}/fp64_sim": f_sim,
            f"{name}/single_idx": single_idx, f"{name}/single_sim": single_sim,
            f"{name}/batch_idx": batch_idx, f"{name}/batch_sim": batch_sim,
            f"{name}/cpu_idx": cpu_idx,
        })
    matrix, query, queries = synth.golden_50k()
    g_idx, g_sim = fp64_topk(matrix, np.vstack([query[None, :], queries]), 5)
    arrays["golden50k/fp64_idx"] = g_idx
    arrays["golden50k/fp64_sim"] = g_sim
    np.savez_compressed(out, **arrays)
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else OUT)
