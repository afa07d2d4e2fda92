// IVF prefilter on the device (ivf.hip).  Mirrors ApproxIndex of
// src/approx.rs:11-25: centroids, a cluster-contiguous reordered copy of the
// rows, the reordered -> original row map and the per-cluster offsets.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

#include "select.hpp"

namespace nndex {

class GpuIndex;

// approx.rs constants and heuristics (MIN_ROWS_FOR_IVF :31, KMEANS_MAX_ITERS
// :28, choose_nlist :305-311, choose_nprobe :318-336, isqrt :339-351).
constexpr size_t kMinRowsForIvf = 256;
constexpr uint32_t kKmeansMaxIters = 12;
size_t ivf_isqrt(size_t n);
size_t ivf_choose_nlist(size_t rows);
size_t ivf_choose_nprobe(size_t nlist, size_t k, size_t rows);

struct IvfIndex {
    uint32_t rows = 0, pdims = 0, nlist = 0, iters = 0;
    float *centroids = nullptr;   // [nlist][pdims] row-major
    uint32_t *offsets = nullptr;  // [nlist + 1]
    uint32_t *orig = nullptr;     // [rows] original row of each reordered position
    uint32_t *assign = nullptr;   // [rows] final cluster of each original row
    float4 *xr = nullptr;         // reordered rows, tiled FP32
    uint32_t rechecked = 0;       // rows re-evaluated exactly, summed over iterations
    double build_ms = 0.0;
    std::unique_ptr<GpuIndex> cindex;  // centroid index used for probing
    // host: prefix sums of the list sizes sorted descending, so
    // max_probe_rows(n) bounds the rows of any n probed lists
    std::vector<uint64_t> size_prefix;
    size_t max_probe_rows(size_t nprobe) const {
        if (size_prefix.empty()) return rows;
        return size_prefix[nprobe < size_prefix.size() ? nprobe : size_prefix.size() - 1];
    }
    // fills size_prefix from the device offsets (after a build or load)
    void read_sizes(hipStream_t s);
    // ApproxIndex::provides_speedup (approx.rs:34-36)
    bool provides_speedup() const { return rows >= kMinRowsForIvf && nlist >= 4; }
    ~IvfIndex();
};

// List-major scan of the probed lists.  The (query, probe) pairs are sorted by
// list; each wavefront takes up to QB pairs of one list and one row chunk of
// it, so QB queries share every row load, and writes each pair's best k keys
// for that chunk; a per-query merge gives the final top-k.
struct IvfScanArgs {
    const float4 *x;            // reordered rows, tiled
    uint32_t d4;
    const float *q;             // [nq][d4 * 4] normalised padded queries
    const uint64_t *probe;      // [nq][nprobe] probed cluster ids
    uint32_t nq, nprobe, nlist, k;
    uint32_t kk;                // keys kept per partial list (ivf_select_k(k))
    uint32_t qb;                // pairs per wavefront (1, 4 or 8)
    uint32_t nchunk;            // row chunks per list
    const uint32_t *offsets;    // [nlist + 1] list offsets in the reordered rows
    const uint32_t *orig;       // [rows] original row of each reordered position
    // scratch (IvfScratch::reserve sizes them)
    uint32_t *pair_key, *pair_val, *sorted_key, *sorted_pair;  // [nq * nprobe]
    uint32_t *pair_base;        // [nq * nprobe] rows of the query's earlier probed lists
    uint32_t *pair_cnt, *pair_off;  // [nlist], [nlist + 1]
    uint32_t *grp_cnt, *grp_off;    // [nlist], [nlist + 1]
    uint64_t *part;             // [nq * nprobe][nchunk][kk]
    void *sort_tmp;
    size_t sort_bytes;
    uint64_t *out_idx;          // [nq][k]
    float *out_sim;             // [nq][k]
    uint64_t *out_len;          // [nq]
    // large-k form: every probed row's key (score, arrival index) is written
    // to keys_out[q][arrival] (no per-list top-k, no merge) and each query's
    // probed row count to totals[q]
    uint64_t *keys_out;
    size_t keys_ld;
    uint32_t *totals;           // [nq] (may be null outside the large-k form)
};

// Whether k results need the large-k form: the list scan's LDS lists keep
// 2k - 1 keys (the reference heap's tie rule) only up to 960.
inline bool ivf_large_k(size_t k) { return 2 * k - 1 > 960; }

// Large-k results from the selection over keys_out (b.sorted holds each
// query's top min(2k - 1, totals[q]) keys): the reference heap's choice among
// rows tied at the k-th score over arrival order, the lowest original rows
// evicted first, results in (score desc, original row asc) order, and
// out_len[q] = min(k, totals[q]).  Uses b.cand and b.cnt as scratch.
void ivf_finish_large_k(const IvfScanArgs &a, const SelectBuffers &b, uint32_t k,
                        hipStream_t s);

// Keys the list scan keeps per (query, list, chunk) for k results: 2k - 1, so
// the merge can apply the reference heap's choice among tied rows (k <= 480;
// larger k take the large-k form).
uint32_t ivf_select_k(size_t k);
// Launch shape of the list scan for a batch: pairs per wavefront and row
// chunks per list (enough wavefronts to fill the device, a bounded merge).
void ivf_scan_shape(size_t nq, size_t nprobe, size_t nlist, size_t rows, size_t k, uint32_t &qb,
                    uint32_t &nchunk);
// Scratch bytes of the list scan for nq queries (worst case over shapes).
size_t ivf_scan_scratch_bytes(size_t nq, size_t nprobe, size_t nlist, size_t rows, size_t k);
// rocPRIM temporary bytes for sorting `pairs` (list, pair) keys of `nlist` lists.
size_t ivf_sort_bytes(size_t pairs, size_t nlist);

// FP8 copy of the rows for the k-means++ pre-test: tiled units of 16 E4M3
// codes, per-row scales and the largest row error norm |s_r x~_r - x_r|.
struct Fp8Rows {
    const uint4 *codes = nullptr;
    const float *scale = nullptr;
    double err = 0.0;
};
// Build the index of `rows` tiled rows (pdims padded columns) on stream s.
// x_max is the largest row norm; above 512 padded dims the assignment runs on
// the BF16 copy xbf16, whose largest row rounding-error norm is x_err; the
// k-means++ passes test rows on the FP8 copy when fp8.codes is set, else on
// the BF16 copy.
void ivf_build(IvfIndex &ivf, const float4 *x, const uint4 *xbf16, uint32_t rows,
               uint32_t pdims, uint32_t nlist, uint32_t iters, double x_max, double x_err,
               const Fp8Rows &fp8, uint32_t ncu, hipStream_t s);
// Install a saved build (host arrays, already validated: offsets[0] = 0,
// non-decreasing, offsets[nlist] = rows; orig a permutation; assign consistent
// with both) and gather the reordered copy from the resident rows x.
void ivf_load(IvfIndex &ivf, const float4 *x, uint32_t rows, uint32_t pdims, uint32_t nlist,
              const float *centroids, const uint32_t *offsets, const uint32_t *orig,
              const uint32_t *assign, hipStream_t s);
// Scan the probed lists and merge per query (probe ids already on device).
void ivf_scan(const IvfScanArgs &a, hipStream_t s);

}  // namespace nndex
