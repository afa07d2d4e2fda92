// Device top-k selection for any k.
//
// The fused scan keeps each wave's list in LDS, which bounds k (the
// reference heap's tie rule needs the canonical top 2k - 1, so k <= 992
// there).  The reference accepts any k <= rows (src/lib.rs:347) and selects
// with a bounded heap on the host (src/gpu.rs:575-589, src/topk.rs:44-152).
// For larger k this file scores every row into a 64-bit ranking key, finds
// the key of rank ksel by a radix select over the keys (passes of 11 bits:
// per-block LDS histograms, then one block per query picks the bin holding
// the rank), compacts the keys at or above it, sorts them with rocPRIM and
// applies the heap's tie rule.
#include <cstring>  // rocPRIM's headers use memset

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include "common.hpp"
#include "select.hpp"

namespace nndex {

// ---- scoring ----------------------------------------------------------------

// Lane r of a 64-row tile walks row r's float4 chunks in order (four FMA
// chains, (l0 + l1) + (l2 + l3): the reference dot order); the queries are
// wave-uniform and come through the scalar cache.
template <int QB>
__global__ __launch_bounds__(256) void score_keys_kernel(const float4 *__restrict__ x,
                                                         uint32_t rows, uint32_t d4,
                                                         const float *__restrict__ q, uint32_t nq,
                                                         uint64_t *__restrict__ keys, size_t ld) {
    const uint32_t lane = __lane_id();
    const uint32_t wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t ntiles = (rows + 63) / 64;
    const uint32_t nwaves = gridDim.x * 4;
    const_as<f32x4> *Q = as_constant(reinterpret_cast<const f32x4 *>(q));
    for (uint32_t t = blockIdx.x * 4 + wid; t < ntiles; t += nwaves) {
        const float4 *b = x + (size_t)t * d4 * 64 + lane;
        float acc[QB][4];
#pragma unroll
        for (int j = 0; j < QB; ++j) acc[j][0] = acc[j][1] = acc[j][2] = acc[j][3] = 0.f;
#pragma unroll 4
        for (uint32_t c = 0; c < d4; ++c) {
            const float4 v = load_nt(b + (size_t)c * 64);
#pragma unroll
            for (int j = 0; j < QB; ++j) {
                const uint32_t jj = (uint32_t)j < nq ? (uint32_t)j : nq - 1;
                const f32x4 qv = Q[(size_t)jj * d4 + c];
                acc[j][0] = __builtin_fmaf(v.x, qv.x, acc[j][0]);
                acc[j][1] = __builtin_fmaf(v.y, qv.y, acc[j][1]);
                acc[j][2] = __builtin_fmaf(v.z, qv.z, acc[j][2]);
                acc[j][3] = __builtin_fmaf(v.w, qv.w, acc[j][3]);
            }
        }
        const uint32_t row = t * 64 + lane;
        if (row < rows) {
#pragma unroll
            for (int j = 0; j < QB; ++j)
                if ((uint32_t)j < nq)
                    keys[(size_t)j * ld + row] =
                        make_key((acc[j][0] + acc[j][1]) + (acc[j][2] + acc[j][3]), row);
        }
    }
}

void launch_score_keys(const float4 *x, uint32_t rows, uint32_t d4, const float *q, uint32_t nq,
                       uint64_t *keys, size_t ld, uint32_t ncu, hipStream_t s) {
    if (nq == 0 || nq > 4) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "score_keys takes 1-4 queries");
    if (ld < rows) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "key array shorter than the rows");
    const uint32_t tiles = (rows + 63) / 64;
    uint32_t grid = (tiles + 3) / 4;
    if (grid > ncu * 8) grid = ncu * 8;
    if (grid == 0) grid = 1;
    if (nq == 1)
        hipLaunchKernelGGL(score_keys_kernel<1>, dim3(grid), dim3(256), 0, s, x, rows, d4, q, nq,
                           keys, ld);
    else if (nq == 2)
        hipLaunchKernelGGL(score_keys_kernel<2>, dim3(grid), dim3(256), 0, s, x, rows, d4, q, nq,
                           keys, ld);
    else
        hipLaunchKernelGGL(score_keys_kernel<4>, dim3(grid), dim3(256), 0, s, x, rows, d4, q, nq,
                           keys, ld);
    NNDEX_HIP_CHECK(hipGetLastError());
}

// ---- radix select -------------------------------------------------------------

// Candidates of query q (never past its key array).
__device__ inline uint32_t sel_count(const SelectBuffers &b, uint32_t q) {
    const uint32_t n = b.n ? b.n[q] : b.n_all;
    return n < b.ld ? n : static_cast<uint32_t>(b.ld);
}

// Keys of the pass: those whose bits above `shift` equal the prefix's.
__device__ inline bool sel_match(uint64_t key, const SelState &st) {
    return st.shift >= 64 || (key >> st.shift) == (st.prefix >> st.shift);
}

__global__ __launch_bounds__(256) void select_init_kernel(SelectBuffers b) {
    const uint32_t q = blockIdx.x;
    for (uint32_t i = threadIdx.x; i < kSelectBins; i += 256) b.hist[(size_t)q * kSelectBins + i] = 0;
    if (threadIdx.x == 0) {
        const uint32_t n = sel_count(b, q);
        SelState st;
        st.prefix = 0;
        st.shift = 64;
        st.remaining = b.ksel;
        // every key selected: the threshold 0 admits all of them
        st.done = b.ksel >= n ? 1u : 0u;
        st.pad = 0;
        b.st[q] = st;
        b.cnt[q] = 0;
    }
}

// Histogram of the next (up to) 11 bits of the keys sharing the prefix.
__global__ __launch_bounds__(256) void select_hist_kernel(SelectBuffers b) {
    __shared__ uint32_t h[kSelectBins];
    const uint32_t q = blockIdx.y;
    const SelState st = b.st[q];
    if (st.done) return;  // block-uniform
    const uint32_t bits = st.shift < 11 ? st.shift : 11;
    const uint32_t lo = st.shift - bits, nb = 1u << bits;
    for (uint32_t i = threadIdx.x; i < nb; i += 256) h[i] = 0;
    __syncthreads();
    const uint32_t n = sel_count(b, q);
    const uint64_t *keys = b.keys + (size_t)q * b.ld;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const uint64_t key = keys[i];
        if (sel_match(key, st)) atomicAdd(&h[(uint32_t)(key >> lo) & (nb - 1)], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nb; i += 256)
        if (h[i]) atomicAdd(b.hist + (size_t)q * kSelectBins + i, h[i]);
}

// One block per query: the bin (from the top) holding the rank; the prefix
// grows by that bin, and the selection is done when every key of the bin is
// needed.  Clears the histogram for the next pass.
__global__ __launch_bounds__(256) void select_pick_kernel(SelectBuffers b) {
    __shared__ uint32_t part[256];
    __shared__ uint32_t sel_bin, sel_above;
    const uint32_t q = blockIdx.x, t = threadIdx.x;
    SelState st = b.st[q];
    if (st.done) return;  // block-uniform
    const uint32_t bits = st.shift < 11 ? st.shift : 11;
    const uint32_t nb = 1u << bits;
    uint32_t *hist = b.hist + (size_t)q * kSelectBins;
    // thread t covers descending bins nb-1-i for i in [8t, 8t + 8)
    uint32_t c[8], sum = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint32_t i = 8 * t + j;
        c[j] = i < nb ? hist[nb - 1 - i] : 0u;
        sum += c[j];
    }
    part[t] = sum;
    if (t == 0) {
        sel_bin = 0;
        sel_above = 0;
    }
    __syncthreads();
    for (uint32_t off = 1; off < 256; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    const uint32_t incl = part[t], excl = incl - sum;
    if (excl < st.remaining && incl >= st.remaining) {
        uint32_t run = excl;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (run + c[j] >= st.remaining) {
                sel_bin = nb - 1 - (8 * t + j);
                sel_above = run;
                break;
            }
            run += c[j];
        }
    }
    __syncthreads();
    if (t == 0) {
        const uint32_t bin = sel_bin;
        const uint32_t rem = st.remaining - sel_above;
        const uint32_t in_bin = hist[bin];
        st.shift -= bits;
        st.prefix |= (unsigned long long)bin << st.shift;
        st.remaining = rem;
        st.done = (in_bin == rem || st.shift == 0) ? 1u : 0u;
        b.st[q] = st;
    }
    __syncthreads();
    for (uint32_t i = t; i < nb; i += 256) hist[i] = 0;
}

// Keys at or above the threshold (exactly min(ksel, n) of them: keys are
// unique) into cand[q][...], in any order.
__global__ __launch_bounds__(256) void select_compact_kernel(SelectBuffers b) {
    const uint32_t q = blockIdx.y, lane = __lane_id(), wid = threadIdx.x >> 6;
    const unsigned long long thr = b.st[q].prefix;
    const uint32_t n = sel_count(b, q);
    const uint64_t *keys = b.keys + (size_t)q * b.ld;
    uint64_t *dst = b.cand + (size_t)q * b.ksel;
    for (uint32_t i0 = blockIdx.x * 256 + wid * 64; i0 < n; i0 += gridDim.x * 256) {
        const uint32_t i = i0 + lane;
        const uint64_t key = i < n ? keys[i] : 0;
        const bool pass = i < n && key >= thr;
        const uint64_t m = __ballot(pass);
        if (m == 0) continue;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(b.cnt + q, (uint32_t)__popcll(m));
        base = __shfl(base, 0);
        const uint32_t pos = base + (uint32_t)__popcll(m & __lanemask_lt());
        if (pass && pos < b.ksel) dst[pos] = key;
    }
}

__global__ void select_offsets_kernel(SelectBuffers b) {
    for (uint32_t q = threadIdx.x; q < b.nq; q += blockDim.x) {
        const uint32_t c = b.cnt[q] < b.ksel ? b.cnt[q] : b.ksel;
        b.offs[q] = q * b.ksel;
        b.offs[b.nq + q] = q * b.ksel + c;
    }
}

size_t select_sort_bytes(size_t nq, size_t ksel) {
    size_t a = 0, c = 0;
    NNDEX_HIP_CHECK(rocprim::radix_sort_keys_desc(nullptr, a, (uint64_t *)nullptr,
                                                  (uint64_t *)nullptr, (unsigned)ksel, 0, 64));
    NNDEX_HIP_CHECK(rocprim::segmented_radix_sort_keys_desc(
        nullptr, c, (uint64_t *)nullptr, (uint64_t *)nullptr, (unsigned)(nq * ksel),
        (unsigned)nq, (uint32_t *)nullptr, (uint32_t *)nullptr, 0, 64));
    return (a > c ? a : c) + 256;
}

void select_top(const SelectBuffers &b, hipStream_t s) {
    if (b.nq == 0) return;
    if (b.ksel == 0) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "select of zero keys");
    if ((size_t)b.nq * b.ksel > 0xFFFFFFFFull)
        throw Error(NNDEX_ERR_UNSUPPORTED, "selection too large for one group");
    const size_t nmax = b.n ? b.ld : b.n_all;
    uint32_t bx = (uint32_t)((nmax + 256 * 16 - 1) / (256 * 16));
    if (bx < 1) bx = 1;
    if (bx > 1024) bx = 1024;
    hipLaunchKernelGGL(select_init_kernel, dim3(b.nq), dim3(256), 0, s, b);
    // 64 key bits: five passes of 11 bits and one of 9 (each a no-op once
    // the query's selection is done)
    for (int pass = 0; pass < 6; ++pass) {
        hipLaunchKernelGGL(select_hist_kernel, dim3(bx, b.nq), dim3(256), 0, s, b);
        hipLaunchKernelGGL(select_pick_kernel, dim3(b.nq), dim3(256), 0, s, b);
    }
    hipLaunchKernelGGL(select_compact_kernel, dim3(bx, b.nq), dim3(256), 0, s, b);
    NNDEX_HIP_CHECK(hipGetLastError());
    select_sort(b, s);
}

void select_sort(const SelectBuffers &b, hipStream_t s) {
    if (b.nq == 0) return;
    size_t bytes = b.sort_bytes;
    if (!b.n) {
        // every query selects min(ksel, n_all) keys: one device-wide sort each
        const uint32_t c = b.ksel < b.n_all ? b.ksel : b.n_all;
        for (uint32_t q = 0; q < b.nq; ++q)
            NNDEX_HIP_CHECK(rocprim::radix_sort_keys_desc(
                b.sort_tmp, bytes, b.cand + (size_t)q * b.ksel, b.sorted + (size_t)q * b.ksel, c,
                0, 64, s));
    } else {
        hipLaunchKernelGGL(select_offsets_kernel, dim3(1), dim3(256), 0, s, b);
        NNDEX_HIP_CHECK(rocprim::segmented_radix_sort_keys_desc(
            b.sort_tmp, bytes, b.cand, b.sorted, b.nq * b.ksel, b.nq, b.offs, b.offs + b.nq, 0,
            64, s));
    }
    NNDEX_HIP_CHECK(hipGetLastError());
}

// ---- the reference heap's tie rule over a sorted canonical list -------------

// One block per query.  With s the k-th score: if rows tie at s across the
// k-th place, the reference's heap (rows pushed in row order, the lowest row
// evicted first among equal minima) keeps G (rows above s) and the tied rows
// up to t*, the k-th smallest row of those >= s, less the e lowest of them,
// e = |G rows after t*|; so result i is L[i] for i < |G| and L[i + e] after.
__global__ __launch_bounds__(1024) void finish_exact_kernel(SelectBuffers b, uint32_t k,
                                                            uint32_t heap_ties,
                                                            uint64_t *__restrict__ out_idx,
                                                            float *__restrict__ out_sim,
                                                            size_t out_stride) {
    __shared__ uint32_t h[256];
    __shared__ uint32_t so[2];
    __shared__ uint32_t g_s, u_s, e_s, tie_s;
    const uint32_t q = blockIdx.x;
    const uint64_t *L = b.sorted + (size_t)q * b.ksel;
    const uint32_t nsel = b.cnt[q] < b.ksel ? b.cnt[q] : b.ksel;
    const uint32_t m = k < nsel ? k : nsel;
    if (threadIdx.x == 0) {
        g_s = 0;
        u_s = 0;
        e_s = 0;
        tie_s = 0;
        if (heap_ties && k >= 2 && k < nsel &&
            (uint32_t)(L[k] >> 32) == (uint32_t)(L[k - 1] >> 32)) {
            const uint32_t sh = (uint32_t)(L[k - 1] >> 32);
            // L is descending: g = first index scoring <= s, u = first < s
            uint32_t lo = 0, hi = nsel;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) / 2;
                if ((uint32_t)(L[mid] >> 32) > sh) lo = mid + 1; else hi = mid;
            }
            g_s = lo;
            lo = 0;
            hi = nsel;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) / 2;
                if ((uint32_t)(L[mid] >> 32) >= sh) lo = mid + 1; else hi = mid;
            }
            u_s = lo;
            tie_s = 1;
        }
    }
    __syncthreads();
    const uint32_t g = g_s;
    if (tie_s) {
        const uint32_t tstar =
            block_select_small(u_s, k, [&](uint32_t i) { return key_row(L[i]); }, h, so);
        if (threadIdx.x == 0) e_s = 0;
        __syncthreads();
        uint32_t local = 0;
        for (uint32_t i = threadIdx.x; i < g; i += blockDim.x) local += key_row(L[i]) > tstar;
        if (local) atomicAdd(&e_s, local);
        __syncthreads();
    }
    const uint32_t e = e_s;
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
        const uint64_t key = L[i < g ? i : i + e];
        out_idx[(size_t)q * out_stride + i] = key_row(key);
        out_sim[(size_t)q * out_stride + i] = key_sim(key);
    }
}

void launch_finish_exact(const SelectBuffers &b, uint32_t k, bool heap_ties, uint64_t *out_idx,
                         float *out_sim, size_t out_stride, hipStream_t s) {
    if (b.nq == 0) return;
    hipLaunchKernelGGL(finish_exact_kernel, dim3(b.nq), dim3(1024), 0, s, b, k,
                       heap_ties ? 1u : 0u, out_idx, out_sim, out_stride);
    NNDEX_HIP_CHECK(hipGetLastError());
}

}  // namespace nndex
