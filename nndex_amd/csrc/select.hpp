// Device top-k selection for any k (select.hip): the path for k beyond the
// fused scan's LDS lists.  Replaces, for large k, the host TopKAccumulator of
// the reference GPU path (src/gpu.rs:575-589 with src/topk.rs:44-152), which
// accepts any k <= rows (src/lib.rs:347).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace nndex {

// Radix-select state of one query: the selected keys are those >= `prefix`
// once `done` is set.
struct SelState {
    unsigned long long prefix;
    uint32_t shift;      // low bits of the key still undecided
    uint32_t remaining;  // rank of the threshold key among keys sharing the prefix
    uint32_t done;
    uint32_t pad;
};

// Buffers of one selection over a group of nq queries, each with its own
// array of candidate keys (unique 64-bit ranking keys, make_key order).
struct SelectBuffers {
    uint64_t *keys = nullptr;     // [nq][ld] candidate keys
    size_t ld = 0;
    const uint32_t *n = nullptr;  // [nq] candidates per query, or nullptr: n_all each
    uint32_t n_all = 0;
    uint32_t nq = 0;
    uint32_t ksel = 0;            // keys selected per query (the top ksel, or all)
    SelState *st = nullptr;       // [nq]
    uint32_t *hist = nullptr;     // [nq][2048], zero between selections
    uint64_t *cand = nullptr;     // [nq][ksel] selected keys, unsorted
    uint64_t *sorted = nullptr;   // [nq][ksel] selected keys, descending
    uint32_t *cnt = nullptr;      // [nq] keys selected
    uint32_t *offs = nullptr;     // [2 nq] segment begins, then ends
    void *sort_tmp = nullptr;
    size_t sort_bytes = 0;
};

// Histogram bins of a radix pass.
constexpr uint32_t kSelectBins = 2048;

// Rank-th smallest (1-based) of the 32-bit values val(i), i < n, over the
// block: four passes of 8 bits with an LDS histogram.  Every thread gets it.
template <typename Val>
__device__ uint32_t block_select_small(uint32_t n, uint32_t rank, Val val, uint32_t *h,
                                       uint32_t *shared_out) {
    uint32_t prefix = 0, remaining = rank;
    for (int pass = 0; pass < 4; ++pass) {
        const uint32_t shift = 24 - 8 * pass;
        for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) h[i] = 0;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
            const uint32_t v = val(i);
            if (pass == 0 || (v >> (shift + 8)) == (prefix >> (shift + 8)))
                atomicAdd(&h[(v >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t cum = 0, bin = 255;
            for (uint32_t j = 0; j < 256; ++j) {
                if (cum + h[j] >= remaining) {
                    bin = j;
                    break;
                }
                cum += h[j];
            }
            shared_out[0] = prefix | (bin << shift);
            shared_out[1] = remaining - cum;
        }
        __syncthreads();
        prefix = shared_out[0];
        remaining = shared_out[1];
        __syncthreads();
    }
    return prefix;
}

// rocPRIM scratch for sorting nq segments of up to ksel keys.
size_t select_sort_bytes(size_t nq, size_t ksel);

// Exact scores of up to 4 queries (row-major, d4 float4 each, the first `nq`
// valid) against every row of the tiled matrix, in the reference dot order,
// as keys[q][row] = make_key(score, row).
void launch_score_keys(const float4 *x, uint32_t rows, uint32_t d4, const float *q, uint32_t nq,
                       uint64_t *keys, size_t ld, uint32_t ncu, hipStream_t s);

// Select the top min(ksel, n_q) keys of every query and sort them
// descending into b.sorted (b.cnt[q] of them).
void select_top(const SelectBuffers &b, hipStream_t s);

// The sort stage alone: cand[q][0, min(cnt[q], ksel)) -> sorted, descending.
void select_sort(const SelectBuffers &b, hipStream_t s);

// Exact-search results from sorted canonical lists: k results per query;
// with `heap_ties` (each list then holds the top min(2k - 1, rows) keys) the
// reference heap's choice among rows tied at the k-th score, else the first k.
void launch_finish_exact(const SelectBuffers &b, uint32_t k, bool heap_ties, uint64_t *out_idx,
                         float *out_sim, size_t out_stride, hipStream_t s);

}  // namespace nndex
