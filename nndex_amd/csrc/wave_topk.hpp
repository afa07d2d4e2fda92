// Wavefront-level bounded top-k over 64-bit ranking keys, kept in LDS.
//
// Replaces the host TopKAccumulator loop of the reference GPU path
// (src/gpu.rs:575-589, :758-776; heap semantics src/topk.rs:44-152) with an
// on-device selection under the total order (similarity desc, row asc).
//
// A wave owns `cap` key slots (cap a power of two, cap >= k + 64).  Candidates
// above the wave's threshold are appended with a ballot; when the next
// append could overflow, the wave bitonic-sorts its slots (descending), keeps
// the first k and raises its threshold to the k-th key.  An optional
// device-wide threshold (atomic max of every wave's k-th key) is folded in:
// no key below it can be in the global top-k.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace nndex {

__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
}

struct WaveTopK {
    uint64_t *buf;    // LDS slots
    uint32_t cap;     // power of two
    uint32_t k;       // keys to keep
    uint32_t rank;    // keys scoring below the rank-th key are dropped (rank <= k)
    uint32_t count;   // wave-uniform
    uint64_t thresh;  // wave-uniform: keys must be > thresh
    unsigned long long *global_thresh;  // device-wide k-th key, or nullptr

    // rank < k keeps, besides the best `rank` keys, up to k - rank more keys
    // that tie the rank-th key's score (lowest rows first): the keys the
    // reference heap's tie rule can need.
    __device__ void init(uint64_t *b, uint32_t c, uint32_t kk, unsigned long long *g,
                         uint32_t rk = 0) {
        buf = b;
        cap = c;
        k = kk;
        rank = rk && rk < kk ? rk : kk;
        count = 0;
        thresh = 0;
        global_thresh = g;
    }

    // Bitonic sort (descending) of the first n slots, n the smallest power of
    // two holding the count (slots past the count are cleared to key 0).
    __device__ void sort_desc() {
        const uint32_t lane = __lane_id();
        uint32_t n = 2;
        while (n < count) n <<= 1;
        if (n > cap) n = cap;
        for (uint32_t i = count + lane; i < n; i += 64) buf[i] = 0;
        wave_lds_sync();
        for (uint32_t size = 2; size <= n; size <<= 1) {
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                for (uint32_t p = lane; p < n / 2; p += 64) {
                    uint32_t lo = 2 * p - (p & (stride - 1));
                    uint32_t hi = lo + stride;
                    bool desc = (lo & size) == 0;
                    uint64_t a = buf[lo], b = buf[hi];
                    if ((a < b) == desc) {
                        buf[lo] = b;
                        buf[hi] = a;
                    }
                }
                wave_lds_sync();
            }
        }
    }

    // Sort, keep the best k (and only keys scoring at least the rank-th
    // key's score), raise the threshold.
    __device__ void compact() {
        sort_desc();
        if (count > k) count = k;
        if (rank < k && count >= rank) {
            // floor: below every key with the rank-th key's score
            const uint64_t floor = (buf[rank - 1] & 0xFFFFFFFF00000000ull) - 1;
            for (uint32_t i0 = rank; i0 < count; i0 += 64) {
                const uint32_t i = i0 + __lane_id();
                const uint64_t m = __ballot(i < count && buf[i] <= floor);
                if (m) {
                    count = i0 + __ffsll(static_cast<unsigned long long>(m)) - 1;
                    break;
                }
            }
            if (floor > thresh) thresh = floor;
        }
        if (k > 0 && count == k) {
            uint64_t kth = buf[k - 1];
            if (kth > thresh) thresh = kth;
            if (global_thresh) {
                unsigned long long old = 0;
                if (__lane_id() == 0)
                    old = __hip_atomic_fetch_max(global_thresh, (unsigned long long)kth,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                old = __shfl(old, 0);
                uint64_t g = old > kth ? old : kth;
                if (g - 1 > thresh) thresh = g - 1;
            }
        }
    }

    // Each lane offers at most one key (valid == false offers nothing).
    __device__ void offer(uint64_t key, bool valid) {
        bool pass = valid && key > thresh;
        uint64_t m = __ballot(pass);
        if (m == 0) return;
        uint32_t n = __popcll(m);
        if (count + n > cap) {
            compact();
            pass = valid && key > thresh;
            m = __ballot(pass);
            if (m == 0) return;
            n = __popcll(m);
        }
        uint32_t pos = count + __popcll(m & __lanemask_lt());
        if (pass) buf[pos] = key;
        count += n;
        wave_lds_sync();
    }
};

// The reference GPU path pushes every row, in row order, into one bounded
// heap (src/gpu.rs:575-589 with src/topk.rs:70-105): a full heap rejects a
// score <= its minimum and otherwise evicts its smallest item, the lowest
// score with the LOWEST row first.  Where rows tie at the k-th score s this
// keeps a different set than the (score desc, row asc) order: with G the rows
// above s and t* the row at which the k-th row >= s arrives, the heap holds G
// and the tied rows up to t*, less the e lowest of those, e = |G rows after
// t*| (each such row arrives at a full heap whose minimum is a tied row).
//
// Given the canonical top list L[0..n) (keys, descending), returns g (rows
// above s) and e, so that the reference's k_out results are L[i] for i < g
// and L[i + e] for g <= i < k_out.  The tied rows in L are in row order (key
// order); `complete` says L holds G and the k_out lowest tied rows (or every
// tied row): true for a canonical top-(2 k_out - 1) selection, possibly
// trimmed of keys scoring below s (WaveTopK's rank), or for one covering
// every row.  Otherwise e = 0 (canonical order).  Wave-uniform.
// Ties are equal ordered bits, so +0 and -0 at the boundary count as distinct
// (the reference's float test treats them as equal).
__device__ inline void heap_tie_shift(const uint64_t *L, uint32_t n, uint32_t k_out,
                                      bool complete, uint32_t &g_out, uint32_t &e_out) {
    const uint32_t lane = __lane_id();
    g_out = 0;
    e_out = 0;
    if (k_out < 2 || n <= k_out || !complete) return;
    const uint32_t hi = static_cast<uint32_t>(L[k_out - 1] >> 32);
    if (static_cast<uint32_t>(L[k_out] >> 32) != hi) return;  // no tie across the boundary
    uint32_t g = 0, u = 0;
    for (uint32_t i0 = 0; i0 < n; i0 += 64) {
        const uint32_t i = i0 + lane;
        const uint32_t h = i < n ? static_cast<uint32_t>(L[i] >> 32) : 0;
        g += __popcll(__ballot(i < n && h > hi));
        u += __popcll(__ballot(i < n && h >= hi));
    }
    // t*: the k_out-th smallest row among the u rows >= s
    uint32_t tstar = 0xFFFFFFFFu;
    for (uint32_t i0 = 0; i0 < u; i0 += 64) {
        const uint32_t i = i0 + lane;
        bool hit = false;
        uint32_t r = 0;
        if (i < u) {
            r = key_row(L[i]);
            uint32_t rank = 0;
            for (uint32_t j = 0; j < u; ++j) rank += key_row(L[j]) < r;
            hit = rank == k_out - 1;
        }
        const uint64_t m = __ballot(hit);
        if (m) {
            tstar = __shfl(r, __ffsll(static_cast<unsigned long long>(m)) - 1);
            break;
        }
    }
    uint32_t e = 0;
    for (uint32_t i0 = 0; i0 < g; i0 += 64) {
        const uint32_t i = i0 + lane;
        e += __popcll(__ballot(i < g && key_row(L[i]) > tstar));
    }
    g_out = g;
    e_out = e;
}

}  // namespace nndex
