"""Host-side replay of the IVF search's index arithmetic (no GPU): the pair
grouping of the list scan, each wavefront's row chunk, the arrival index of
every probed row, the key slot the large-k form writes, the partial-list slot
of the merge form and the merge's arrival -> original row lookup, each
checked against the length of the buffer it addresses.

The data are the oracle's own IVF build and probes (orc.Ivf) at the shapes
the GPU tests run, including the large-k case that faulted the GPU (the key slot was formed as size_t + uint32 + uint32, so a 32-bit
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
    # work items as the kernel decodes them (four waves per block, the grid
    # rounded up): every (group, chunk) exactly once
    seen = set()
    for item in range(-(-(gmax * nchunk) // 4) * 4):
        g, chunk = item // nchunk, item % nchunk
        if g >= grp_off[-1]:
            continue
        c = int(np.searchsorted(grp_off, g, side="right") - 1)
        assert grp_off[c] <= g < grp_off[c + 1]
        assert (g, chunk) not in seen
        seen.add((g, chunk))
        p0 = int(pair_off[c] + (g - grp_off[c]) * qb)
        pn = min(int(pair_off[c + 1]) - p0, qb)
        assert pn >= 1
        pids = order[p0:p0 + pn]
        l0, ln = int(offsets[c]), int(sizes[c])
        per = -(-ln // nchunk)
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
    assert len(seen) == int(grp_off[-1]) * nchunk
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
    (8000, 64, 20, 3, 12),     # the merge-form case kept off the GPU (unexplained fault)
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


# ---------------------------------------------------------------------------
# The MFMA scan's ring of LDS query buffers (nndex_amd/csrc/batch.hip,
# mfma_scan_kernel with DMA staging): per block tile, a prologue copies
# chunks ch0 .. ch0 + NB - 2 into buffers 0 .. NB - 2 and waits for them;
# iteration j (chunk ch0 + j, buffer j % NB) then issues chunk j + NB - 1
# into buffer (j + NB - 1) % NB, waits until at most (ahead - 1) * my_cnt of
# its copies are in flight, and meets the others at a barrier.  The replay
# keeps every wave's in-flight copies in issue order, lets a copy land only
# when a wait forces it (the latest it may), and checks that a buffer is
# never read before its chunk has landed, nor written while another chunk
# is read from it; and that each copy stays inside the image, threshold and
# scale arrays and the LDS ring.
# ---------------------------------------------------------------------------

def _nbuf(maxs, nt, scaled, nw=8, queue=True):
    buf = (maxs * 4 * 16 * nt + 4 * nt + (4 * nt if scaled else 0)) * 16
    q = nw * 128 * 12 if queue else 0
    lds = 160 * 1024
    return 4 if 4 * buf + q <= lds else (3 if 3 * buf + q <= lds else 2)


def ring_replay(nq, maxs, nt, mode, scaled, nw=8, qsplit=1, tiles=2, midbar=False):
    qc = 16 * nt
    ru = 4 * maxs
    q_units = qc * ru
    tu, su = qc // 4, (qc // 4 if scaled else 0)
    buf_units = q_units + tu + su
    am = mode == 2
    nb = _nbuf(maxs, nt, scaled, nw, queue=not am)
    assert (nb * buf_units + (0 if am else nw * 128 * 12 // 16)) * 16 <= 160 * 1024
    ni = nt * maxs
    nch_all = -(-nq // qc)
    image_units = nch_all * qc * ru
    assert image_units <= (nq + 64) * 128      # mfma_image_units
    thresh_len = nq + 64                        # Workspace::reserve_batch
    stage_th = mode == 0

    def copies(w, ch, buf):
        """The copies wave w issues for chunk ch: (lds lo, lds hi, src lo, src hi, array)."""
        out = []
        for i in range(w, ni, nw):
            out.append((buf * buf_units + i * 64, buf * buf_units + i * 64 + 64,
                        ch * q_units + i * 64, ch * q_units + i * 64 + 64, "img"))
        if stage_th and w == ni % nw:
            out.append((buf * buf_units + q_units, buf * buf_units + q_units + tu,
                        ch * qc, ch * qc + qc, "thr"))
        if scaled and w == (ni + 1) % nw:
            out.append((buf * buf_units + q_units + tu, buf * buf_units + q_units + tu + su,
                        ch * qc, ch * qc + qc, "scale"))
        for lo, hi, s0, s1, arr in out:
            assert 0 <= lo < hi <= nb * buf_units
            assert hi - lo <= 64  # one copy instruction: at most 64 lanes x 16 bytes
            if arr == "img":
                assert 0 <= s0 and s1 <= image_units
            else:
                assert s1 <= thresh_len
        return out

    for split in range(qsplit):
        per = -(-nch_all // qsplit)
        ch0, ch1 = split * per, min(split * per + per, nch_all)
        if ch0 >= ch1:
            continue
        for _ in range(tiles):
            inflight = [[] for _ in range(nw)]   # per wave: (buf, chunk) in issue order
            landed = {}                          # buf -> chunk whose copies all landed
            pending = {}                         # buf -> set of (wave, chunk) in flight

            def issue(w, ch, buf):
                cnt = 0
                for c in copies(w, ch, buf):
                    inflight[w].append((buf, ch))
                    pending.setdefault(buf, set()).add((w, ch))
                    cnt += 1
                return cnt

            def wait(w, n):
                while len(inflight[w]) > n:
                    buf, ch = inflight[w].pop(0)
                    if all(x != (buf, ch) for x in inflight[w]):
                        pending[buf].discard((w, ch))

            my_cnt = [len(copies(w, ch0, 0)) for w in range(nw)]
            for j in range(nb - 1):
                if ch0 + j < ch1:
                    for w in range(nw):
                        issue(w, ch0 + j, j)
            for w in range(nw):
                wait(w, 0)
            for j in range(nb - 1):
                if ch0 + j < ch1:
                    assert not pending.get(j)
                    landed[j] = ch0 + j
            if midbar:
                _mid_chunks(ch0, ch1, nb, nw, my_cnt, issue, wait, inflight, pending, landed)
                continue
            for ch in range(ch0, ch1):
                jj = ch - ch0
                cur = jj % nb
                # the barrier before this iteration: its chunk landed, and no
                # copy into its buffer is in flight
                assert landed.get(cur) == ch, (ch, cur, landed)
                assert not pending.get(cur), (ch, cur, pending.get(cur))
                for w in range(nw):
                    if ch + nb - 1 < ch1:
                        tgt = (jj + nb - 1) % nb
                        assert tgt != cur
                        landed.pop(tgt, None)
                        assert issue(w, ch + nb - 1, tgt) == my_cnt[w]
                    ahead = nb - 1 if ch + nb - 1 < ch1 else ch1 - 1 - ch
                    wait(w, (ahead - 1) * my_cnt[w] if ahead >= 2 else 0)
                # barrier: the next chunk's copies have landed in every wave
                nxt = ch + 1
                if nxt < ch1:
                    b = (jj + 1) % nb
                    assert not pending.get(b), (nxt, b, pending.get(b))
                    landed[b] = nxt
            assert all(not inflight[w] for w in range(nw))


def _mid_chunks(ch0, ch1, nb, nw, my_cnt, issue, wait, inflight, pending, landed):
    """The BF16 scans' mid-chunk schedule (three or more buffers): chunk ch's
    barrier sits after step group 0 in waves 0 .. nw/2 - 1 and after the last
    group in the others; before it a wave waits until chunk ch + 1's copies
    have landed (at most (ahead_mid - 1) * my_cnt of its copies in flight,
    ahead_mid = min(nb - 2, ch1 - 1 - ch)), after it the wave copies chunk
    ch + nb - 1 into chunk ch - 1's buffer.  Between the barriers of chunks
    ch - 1 and ch the waves, in any order, read the rest of chunk ch - 1 and
    the first part of chunk ch and issue chunk ch + nb - 2's copies: the
    buffers read and the buffers written or still receiving copies in that
    span must be disjoint, and chunk ch must have landed before it."""
    assert nb >= 3
    for ch in range(ch0, ch1):
        jj = ch - ch0
        cur = jj % nb
        # span from the previous barrier (or the prologue's) to this one
        reads = {cur} | ({(jj - 1) % nb} if ch > ch0 else set())
        assert landed.get(cur) == ch, (ch, cur, landed)
        if ch > ch0:
            prev = ch - 1
            if prev + nb - 1 < ch1:
                tgt = (jj - 1 + nb - 1) % nb
                assert tgt not in reads, (ch, tgt, reads)
                landed.pop(tgt, None)
                for w in range(nw):
                    assert issue(w, prev + nb - 1, tgt) == my_cnt[w]
        # copies in flight during the span land only in buffers not read in it
        for buf, waves in pending.items():
            if waves:
                assert buf not in reads, (ch, buf, reads)
        # this chunk's barrier: every wave waits for chunk ch + 1's copies
        ahead_mid = min(nb - 2, ch1 - 1 - ch)
        for w in range(nw):
            wait(w, (ahead_mid - 1) * my_cnt[w] if ahead_mid >= 2 else 0)
        nxt = ch + 1
        if nxt < ch1:
            b = (jj + 1) % nb
            assert not pending.get(b), (nxt, b, pending.get(b))
            landed[b] = nxt
    # after the last chunk's barrier: the rest of it is read and no copy is
    # issued (chunk ch1 - 1 + nb - 1 is past the range); nothing in flight
    assert all(not inflight[w] for w in range(nw))


@pytest.mark.parametrize("nq,maxs,nt,mode,qsplit", [
    (2048, 12, 2, 0, 1),   # BF16 384 dims, 32-query chunks: 4 buffers
    (2048, 12, 2, 1, 2),   # its sample scan, queries split over 2 block rows
    (2048, 12, 4, 0, 1),   # 64-query chunks: 3 buffers
    (2048, 12, 4, 1, 3),
    (1024, 24, 1, 0, 1),   # BF16 768 dims, 16-query chunks: 4 buffers
    (1024, 24, 1, 1, 5),
    (1024, 24, 2, 0, 1),   # 32-query chunks: 3 buffers
    (17, 4, 4, 0, 1),      # one partial chunk
    (100, 8, 2, 0, 1),     # tail shorter than the ring
    (130, 12, 4, 0, 1),    # three chunks over three buffers
    (4096, 16, 2, 1, 8),
    (4096, 32, 1, 0, 1),   # BF16 1024 dims
])
def test_mfma_scan_mid_chunk_barrier_schedule(nq, maxs, nt, mode, qsplit):
    if _nbuf(maxs, nt, False) < 3:
        pytest.skip("two-buffer rings keep the end-of-chunk barrier")
    ring_replay(nq, maxs, nt, mode, False, qsplit=qsplit, midbar=True)


@pytest.mark.parametrize("nq,maxs,nt,mode,scaled,qsplit", [
    (2048, 12, 2, 0, False, 1),   # BF16 384 dims: 4 buffers
    (2048, 12, 2, 1, False, 2),   # its sample scan, queries split over 2 block rows
    (1024, 24, 2, 0, False, 1),   # BF16 768 dims: 3 buffers
    (1024, 24, 1, 0, False, 1),
    (1024, 12, 2, 0, True, 1),    # FP8 768 dims: thresholds and scales
    (1024, 32, 1, 1, True, 3),
    (17, 4, 4, 0, False, 1),      # one partial chunk
    (100, 8, 2, 0, False, 1),     # tail shorter than the ring
    (33, 2, 4, 1, False, 1),
    (4096, 32, 1, 0, False, 8),
    (4096, 8, 2, 2, False, 1),    # row argmax over 4,096 centroids at 128 dims
    (89, 4, 4, 2, False, 1),      # row argmax, 89 centroids at 64 dims
    (300, 12, 4, 2, False, 1),
])
def test_mfma_scan_ring_schedule(nq, maxs, nt, mode, scaled, qsplit):
    ring_replay(nq, maxs, nt, mode, scaled, qsplit=qsplit)


def test_chunk_image_inverts_the_lds_swizzle():
    """chunk_image_kernel writes query n of unit row c at slot
    (n & ~15) | ((n ^ c) & 15); the scan reads slot (l16 ^ (4 (S & 3) + g)) &
    15 of row 4 S + g for the query of lane group position l16: both must
    name the same query."""
    for qc in (16, 32, 64):
        for c in range(128):
            slots = [(n & ~15) | ((n ^ c) & 15) for n in range(qc)]
            assert sorted(slots) == list(range(qc))
            for slot in range(qc):
                n = (slot & ~15) | ((slot ^ c) & 15)   # the image kernel's inverse
                assert slots[n] == slot
    for s_ in range(32):
        for g in range(4):
            u = 4 * s_ + g
            for t in range(4):
                for l16 in range(16):
                    n = t * 16 + l16
                    read_slot = t * 16 + ((l16 ^ (4 * (s_ & 3) + g)) & 15)
                    assert read_slot == (n & ~15) | ((n ^ u) & 15)


def tail_split_choice(blocks, cap, nch):
    """gpu_index.hip grid_for: the split of the last partial round's tiles."""
    best, split = None, 1
    if blocks > cap and blocks % cap:
        full, rem = blocks // cap, blocks % cap
        best = full + 1.0
        for q in range(2, min(8, nch) + 1):
            t = full + -(-rem * q // cap) / q + 0.02 * q
            if t < best - 1e-9:
                best, split = t, q
    return split


@pytest.mark.parametrize("rows,block_rows,nq,qc,cap", [
    (1_000_000, 256, 2048, 32, 256),   # 3,907 tiles: 67 in the last round
    (5_000_000, 128, 1024, 32, 256),
    (5_000_000, 256, 1024, 32, 256),
    (70_000, 128, 40, 16, 256),        # fewer tiles than blocks: no tail split
    (300_000, 256, 17, 16, 256),       # one chunk: nothing to split
    (2_000_000, 128, 333, 16, 256),
])
def test_mfma_work_items_cover_every_tile_and_chunk_once(rows, block_rows, nq, qc, cap):
    """mfma_scan_kernel's work items: the full rounds' tiles take every chunk,
    the last round's tiles are split into tail_split chunk ranges; every
    (row tile, chunk) pair is scanned exactly once and no block runs more
    than the rounds the host planned."""
    rows_padded = -(-rows // 64) * 64
    nblk = -(-rows_padded // block_rows)
    nch = -(-nq // qc)
    grid = min(nblk, cap)
    ts = tail_split_choice(nblk, cap, nch)
    nfull = nblk // grid * grid if ts > 1 else nblk
    nwork = nfull + (nblk - nfull) * ts
    per = -(-nch // ts)
    seen = np.zeros((nblk, nch), np.int32)
    busy = np.zeros(grid)
    for b in range(grid):
        for w in range(b, nwork, grid):
            if w < nfull:
                bt, c0, c1 = w, 0, nch
            else:
                r = w - nfull
                bt, c0 = nfull + r // ts, (r % ts) * per
                c1 = min(c0 + per, nch)
                if c0 >= c1:
                    continue
            seen[bt, c0:c1] += 1
            busy[b] += (c1 - c0) / nch
    assert np.all(seen == 1)
    # the busiest block does no more tile-rounds than the plan
    # (pieces are whole chunks: ceil(nch / tail_split) of nch each)
    plan = nblk // grid + (-(-(nblk % grid) * ts // grid) * per / nch if nblk % grid else 0)
    assert busy.max() <= plan + 1e-9
    if ts > 1:
        assert busy.max() < nblk // grid + 1   # better than a whole extra round
    if (rows, nq) == (1_000_000, 2048):
        assert ts == 3
