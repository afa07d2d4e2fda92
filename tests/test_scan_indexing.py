"""Host-side replay of the IVF search's index arithmetic (no GPU): the pair
grouping of the list scan, each wavefront's row chunk, the arrival index of
every probed row, the key slot the large-k form writes, the partial-list slot
of the merge form and the merge's arrival -> original row lookup, each
checked against the length of the buffer it addresses.

The data are the oracle's own IVF build and probes (orc.Ivf) at the shapes
the GPU tests run, including the case the round's GPU memory fault came from
(large k: the key slot was formed as size_t + uint32 + uint32, so a 32-bit
arrival base that wraps below zero was widened before it could wrap back,
landing 2^32 keys past the array); the pre-fix form is replayed beside the
shipped one and must leave the array.  The arithmetic restates
nndex_amd/csrc/ivf.hip (ivf_scan_shape, pair_base_kernel, group_count_kernel,
ivf_list_scan_kernel, ivf_merge_kernel, ivf_arrival_orig)."""

import numpy as np
import pytest

U32 = 1 << 32


def scan_shape(nq, nprobe, nlist, rows, k):
    """ivf_scan_shape: pairs per wavefront and row chunks per list."""
    pairs = nq * nprobe
    per_list = -(-pairs // nlist)
    qb = 8 if per_list >= 6 else (4 if per_list >= 2 else 1)
    if qb == 8 and k > 192:
        qb = 4
    if qb == 4 and k > 448:
        qb = 1
    waves = -(-pairs // qb)
    nc = -(-16384 // waves)
    by_len = max(rows // nlist // 64, 1)
    by_merge = max(65536 // (nprobe * k), 1)
    return qb, max(min(nc, by_len, by_merge, 64), 1)


def replay(offsets, orig, probes, rows, k, large):
    """Every index the list scan (and merge or large-k form) computes for one
    batch; returns the pre-fix key slots' maximum for the large-k form."""
    nq, nprobe = probes.shape
    nlist = len(offsets) - 1
    sizes = np.diff(offsets).astype(np.int64)
    kk = 1 if large else (2 * k - 1 if 2 * k - 1 <= 960 else k)
    qb, nchunk = scan_shape(nq, nprobe, nlist, rows, 1 if large else kk)
    pairs = nq * nprobe
    flat = probes.reshape(-1).astype(np.int64)
    assert flat.min() >= 0 and flat.max() < nlist
    # pair_base_kernel: arrival base of each (query, probe), the query totals
    base = np.zeros(pairs, np.int64)
    totals = np.zeros(nq, np.int64)
    for q in range(nq):
        acc = 0
        for p in range(nprobe):
            base[q * nprobe + p] = acc
            acc += sizes[probes[q, p]]
        totals[q] = acc
    # large-k key stride: the nprobe largest lists (IvfIndex::max_probe_rows)
    ld = int(np.sort(sizes)[::-1][:nprobe].sum())
    assert totals.max() <= ld
    # the pair sort by list (stable) and the per-list pair groups
    order = np.argsort(flat, kind="stable")
    pair_cnt = np.bincount(flat, minlength=nlist)
    pair_off = np.concatenate([[0], np.cumsum(pair_cnt)])
    grp_cnt = -(-pair_cnt // qb)
    grp_off = np.concatenate([[0], np.cumsum(grp_cnt)])
    gmax = -(-pairs // qb) + min(nlist, pairs)
    assert grp_off[-1] <= gmax  # the launch covers every group
    worst_old = 0
    part_len = pairs * nchunk * kk
    for g in range(int(grp_off[-1])):
        c = int(np.searchsorted(grp_off, g, side="right") - 1)
        assert grp_off[c] <= g < grp_off[c + 1]
        p0 = int(pair_off[c] + (g - grp_off[c]) * qb)
        pn = min(int(pair_off[c + 1]) - p0, qb)
        assert pn >= 1
        pids = order[p0:p0 + pn]
        l0, ln = int(offsets[c]), int(sizes[c])
        per = -(-ln // nchunk)
        for chunk in range(nchunk):
            r0 = l0 + min(chunk * per, ln)
            r1 = l0 + min((chunk + 1) * per, ln)
            row = np.arange(r0 // 64 * 64, r1, dtype=np.int64)
            row = row[(row >= r0) & (row < r1)]
            for pid in pids:
                q = int(pid) // nprobe
                abase = (int(base[pid]) - l0) % U32  # uint32 in the kernel
                arr = (abase + row) % U32             # 32-bit sum: the arrival index
                assert arr.size == 0 or arr.max() < totals[q]
                if large:
                    slot = q * ld + arr                 # shipped form
                    assert slot.size == 0 or slot.max() < nq * ld
                    if row.size:
                        worst_old = max(worst_old, q * ld + abase + int(row.max()))
                else:
                    top = (int(pid) * nchunk + chunk) * kk + kk - 1
                    assert top < part_len
    # the merge / finisher lookup: arrival -> probe -> reordered -> original
    for q in range(nq):
        pb = base[q * nprobe:(q + 1) * nprobe]
        arr = np.arange(totals[q])
        lo = np.searchsorted(pb, arr, side="right") - 1
        idx = offsets[probes[q, lo]].astype(np.int64) + (arr - pb[lo])
        assert idx.min() >= 0 and idx.max() < rows
        got = orig[idx]
        assert len(np.unique(got)) == len(got)
    return worst_old, nq * ld


def _ivf(orc, rows, dims, seed):
    rng = np.random.default_rng(seed)
    centres = rng.normal(size=(max(rows // 200, 4), dims)).astype(np.float32)
    raw = centres[rng.integers(0, centres.shape[0], size=rows)] + \
        0.3 * rng.normal(size=(rows, dims)).astype(np.float32)
    m = orc.normalize_rows(raw, rows, dims)
    return m, rng, orc.Ivf(m)


@pytest.mark.parametrize("rows,dims,k,nprobe,nq", [
    (8000, 64, 20, 3, 12),     # the round-one GPU fault's case (merge form)
    (8000, 64, 20, 3, 1),
    (30000, 32, 481, 20, 7),   # the large-k form whose key slot faulted
    (30000, 32, 3000, 24, 1),
    (6000, 32, 100, 52, 12),
    (4000, 32, 10, 40, 24),    # full 8-query wavefronts
    (9000, 32, 700, 60, 40),   # full 8-query wavefronts in the large-k form
])
def test_ivf_scan_indices_stay_in_bounds(orc, rows, dims, k, nprobe, nq):
    m, rng, ref = _ivf(orc, rows, dims, 3 * rows + dims)
    qs = orc.normalize_rows(rng.normal(size=(nq, dims)).astype(np.float32), nq, dims)
    probes = np.stack([ref.probe(q, nprobe) for q in qs]).astype(np.int64)
    offsets = ref.offsets().astype(np.int64)
    orig = ref.original_indices().astype(np.int64)
    large = 2 * k - 1 > 960
    worst_old, length = replay(offsets, orig, probes, rows, k, large)
    if large:
        # the pre-fix slot leaves the key array (by about 2^32 keys)
        assert worst_old >= length and worst_old >= U32
