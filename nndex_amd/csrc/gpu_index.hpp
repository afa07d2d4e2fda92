// GpuIndex: the device-resident index behind the backend C ABI.
// Replaces `GpuIndex` of src/gpu.rs:93-117.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "batch.hpp"
#include "common.hpp"
#include "ivf.hpp"
#include "kernels.hpp"
#include "select.hpp"

namespace nndex {

struct Neighbor {
    uint64_t index;
    float similarity;
};

// Candidate-scoring precision of the batched path.
enum class Precision : int { F32 = 0, BF16 = 1, FP8 = 2 };

// Scratch of the fused scan: block lists, thresholds and the ticket; plus the
// batched path's buffers, grown on demand.
// Leading dimension of the sample score matrix: rows rounded up to 4, so a
// lane's four consecutive rows are one aligned 16-byte store.
inline size_t sample_ld(size_t rows_s) { return (rows_s + 3) & ~static_cast<size_t>(3); }

struct Workspace {
    int device = 0;
    uint64_t generation = 0;  // bumped whenever a buffer is reallocated
    size_t max_queries = 0;
    size_t max_k = 0;
    uint32_t grid = 0;
    uint64_t *part = nullptr;
    size_t part_elems = 0;
    uint32_t *part_count = nullptr;  // [8][part_elems / 2048]
    unsigned long long *theta = nullptr;  // [8]
    uint32_t *ticket = nullptr;
    // batched MFMA path
    size_t b_queries = 0, b_scores = 0, b_cand = 0;
    uint4 *qlow = nullptr;
    uint4 *qimg = nullptr;  // chunk image of the MFMA scan's queries
    float *delta = nullptr;
    float *scores = nullptr;
    unsigned long long *thresh = nullptr;
    float *thresh_f = nullptr;  // the thresholds' float floors
    float *qscale = nullptr;
    uint64_t *cand = nullptr;
    uint32_t *cand_count = nullptr;
    uint8_t *flags = nullptr;
    void reserve_batch(size_t nq, size_t pdims, size_t rows_s, size_t capq);
    // IVF search: probed cluster ids and scores, the list scan's pair sort
    // and partial keys, and the centroid index's own workspace
    size_t ivf_pairs = 0, ivf_lists = 0, ivf_part_elems = 0, ivf_sort_bytes = 0;
    uint64_t *ivf_probe = nullptr;
    float *ivf_probe_sim = nullptr;
    uint32_t *ivf_pair_u32 = nullptr;  // [4][pairs]: key, val, sorted key, sorted pair
    uint32_t *ivf_list_u32 = nullptr;  // [4][lists + 1]: pair count/offset, group count/offset
    uint64_t *ivf_part = nullptr;
    void *ivf_sort_tmp = nullptr;
    std::unique_ptr<Workspace> sub;
    void reserve_ivf(size_t pairs, size_t nlist, size_t part_elems, size_t sort_bytes);
    // certificate flags of the exact batch path
    uint8_t *xflags = nullptr;
    size_t xflags_n = 0;
    void reserve_xflags(size_t nq);
    // large-k selection (select.hip): per group of nq queries, every row's
    // key [nq][ld], the selected keys [nq][ksel] before and after sorting,
    // the radix-select state, histograms, counts, segment offsets and rocPRIM
    // scratch
    size_t sel_nq = 0, sel_ld = 0, sel_k = 0, sel_tmp_bytes = 0;
    uint64_t *sel_keys = nullptr, *sel_cand = nullptr, *sel_sorted = nullptr;
    SelState *sel_st = nullptr;
    uint32_t *sel_hist = nullptr, *sel_cnt = nullptr, *sel_offs = nullptr;
    uint32_t *sel_n = nullptr;  // [nq] candidates per query (IVF: probed rows)
    void *sel_tmp = nullptr;
    void reserve_select(size_t nq, size_t ld, size_t ksel);
    SelectBuffers select_buffers(uint32_t nq, uint32_t ksel, const uint32_t *n,
                                 uint32_t n_all) const;
    ~Workspace();
};

// Scratch of one fused-scan launch: the blocks' published keys, the live
// counts, the (unused) device-wide thresholds and the block ticket; counts,
// thresholds and ticket are zero between launches (the last block resets
// them).  A Workspace provides it, or the no-allocation entry point's static
// device array.
struct ScanScratch {
    uint64_t *part = nullptr;
    size_t part_elems = 0;
    uint32_t *part_count = nullptr;   // [8]
    unsigned long long *theta = nullptr;  // [8]
    uint32_t *ticket = nullptr;
};

// Sizes of the batched MFMA path for one (k, precision): over-fetch k',
// sample stride and rows, candidate capacity per query, queries per group.
// Exact batches of up to this many queries run the fused scan (one launch);
// larger ones the certified MFMA path, which is faster from five queries on.
constexpr size_t kFusedBatchMax = 4;
// Exact batches over at least this many rows score on BF16 (FP32 below).
constexpr size_t kBf16ExactMinRows = 65536;

struct BatchPlan {
    size_t kp = 0;
    // rank of the filter threshold among the sample scores: about 2 (k' + 1)
    // rows of the index score at or above it (at least 8 sample rows), so
    // the k' + 1 best approximate scores are almost always among the
    // candidates (when fewer pass, the certificate bounds the rest by the
    // threshold itself); each candidate costs the scan an append and a
    // global atomic, so fewer candidates make the scan faster
    size_t jt = 0;
    uint32_t stride = 1;
    size_t rows_s = 0;
    size_t capq = 0;
    size_t group = 4096;
    // device bytes of the workspace buffers for nq queries
    size_t workspace_bytes(size_t nq, size_t pdims) const;
};

// An instantiated HIP graph of one batched search (fixed buffers and shape).
struct BatchGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t generation = 0;  // of the workspace it was captured with
    void launch(hipStream_t s) const;
    ~BatchGraph();
};

// Strided row sample used by the batched path's threshold stage.
struct RowSample {
    uint32_t stride = 0, rows = 0;
    float4 *f32 = nullptr;
    uint4 *bf16 = nullptr;
    uint4 *fp8 = nullptr;
    float *fp8_scale = nullptr;
    unsigned long long *fp8_err = nullptr;
};

struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    void reserve(size_t n);
    ~DeviceBuffer();
};

struct PinnedBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    void reserve(size_t n);
    ~PinnedBuffer();
};

class GpuIndex {
  public:
    // From normalised, padded row-major host storage (src/gpu.rs:138-145).
    static std::unique_ptr<GpuIndex> from_storage(const float *storage, size_t rows,
                                                  size_t padded_dims, bool approx,
                                                  bool enable_cache, int64_t device_index);
    // From raw host rows: upload, then validate/normalise on the device
    // (src/lib.rs:239-241 `ensure_finite + pad_rows` or `normalize_rows`).
    static std::unique_ptr<GpuIndex> from_raw(const float *matrix, size_t rows, size_t dims,
                                              bool normalized, bool approx, bool enable_cache,
                                              int64_t device_index);
    ~GpuIndex();

    size_t rows() const { return rows_; }
    size_t padded_dims() const { return pdims_; }
    int device() const { return device_; }
    const nndex_gpu_device_info &info() const { return info_; }
    const float4 *tiled() const { return x_; }
    uint32_t num_cus() const { return ncu_; }

    // GpuIndex::search / search_batch (src/gpu.rs:321-399): host buffers,
    // normalised padded queries, k already capped.
    std::vector<Neighbor> search(const float *query, size_t k, bool approx);
    std::vector<std::vector<Neighbor>> search_batch(const float *queries, size_t query_rows,
                                                    size_t k, bool approx);
    // Front-end batch path: raw host queries (query_rows x dims) are checked
    // and normalised on the device (normalize_rows, src/lib.rs:461-514) or
    // only checked and padded (ensure_finite + pad_rows) when `normalize` is
    // false, then searched.
    std::vector<std::vector<Neighbor>> search_batch_raw(const float *queries, size_t query_rows,
                                                        size_t dims, bool normalize, size_t k,
                                                        bool approx);
    // Front-end single query (src/lib.rs:335-355): `dims` raw values checked
    // and padded, or normalised in normalize_vector's sequential order
    // (src/lib.rs:558-575), on the device, then searched.
    std::vector<Neighbor> search_raw(const float *query, size_t dims, bool normalize, size_t k,
                                     bool approx);

    // Enqueue an exact search of device-resident queries.  With `heap_ties`,
    // rows tied at the k-th score are chosen as the reference GPU path's
    // bounded heap chooses them (src/gpu.rs:575-589); without, in (score
    // desc, row asc) order.
    void search_device(Workspace &ws, const float *dq, size_t nq, size_t k, uint64_t *d_idx,
                       float *d_sim, hipStream_t s, bool heap_ties = true) const;
    // The same fused scan over caller-provided scratch: each launch takes as
    // many queries (at most 4) as the scratch holds block lists for; throws
    // UNSUPPORTED when not even one query's lists fit.
    void search_device(const ScanScratch &sc, const float *dq, size_t nq, size_t k,
                       uint64_t *d_idx, float *d_sim, hipStream_t s, bool heap_ties = true) const;
    // Block-list keys one single-query fused-scan launch of k results needs.
    size_t scan_part_keys(size_t k, bool heap_ties = true) const;
    // Whether the fused scan's LDS lists hold what k results need (with
    // heap_ties, the canonical top 2k - 1); otherwise the exact search runs
    // the large-k selection.
    bool fused_scan_serves(size_t k, bool heap_ties = true) const;
    // Exact search for any k: every row's key, radix select of the top
    // min(2k - 1, rows), sort, the heap's tie rule (allocates its buffers in
    // the workspace on first use).
    void search_large_k(Workspace &ws, const float *dq, size_t nq, size_t k, uint64_t *d_idx,
                        float *d_sim, hipStream_t s, bool heap_ties = true) const;

    // Enqueue the batched MFMA search with exact rerank and certificate
    // (no host synchronisation).  flags[q] = 1 marks a query whose result is
    // not certified exact; kp = 0 picks the default over-fetch.
    void search_batch_mfma(Workspace &ws, const float *dq, size_t nq, size_t k, Precision p,
                           size_t kp, uint64_t *d_idx, float *d_sim, uint8_t *d_flags,
                           hipStream_t s);
    // Capture search_batch_mfma with these arguments into a graph (one eager
    // run first creates every lazy buffer, so replays allocate nothing).
    std::unique_ptr<BatchGraph> capture_batch(Workspace &ws, const float *dq, size_t nq,
                                              size_t k, Precision p, size_t kp,
                                              uint64_t *d_idx, float *d_sim, uint8_t *d_flags,
                                              hipStream_t s);
    // Exact search of any batch: the fused scan for small batches; for larger
    // ones the F32 MFMA path, then the scan for any uncertified query (this
    // synchronises `s` once to read the flags).
    void search_exact_any(Workspace &ws, const float *dq, size_t nq, size_t k, uint64_t *d_idx,
                          float *d_sim, hipStream_t s, bool heap_ties = true);
    bool mfma_supported(Precision p) const;
    static size_t default_overfetch(size_t k, Precision p);
    BatchPlan plan_batch(size_t k, Precision p, size_t kp) const;
    // Scoring precision of the exact batch path (BF16 up to 1024 padded dims
    // for indexes of kBf16ExactMinRows rows or more, else FP32 up to 512; the
    // certificate and the scan fallback make either exact); false when
    // neither MFMA form applies (fused scan only).
    bool exact_batch_precision(Precision &p) const;
    // Allocate everything a later search of up to nq queries and k results
    // needs (lazy copies, row sample, workspace buffers; IVF index and its
    // centroid workspace when ivf_nprobe > 0), so it allocates nothing.
    void prepare(Workspace &ws, size_t nq, size_t k, size_t ivf_nprobe, hipStream_t s);
    // Device bytes prepare() reserves in a workspace for nq queries, k results.
    size_t workspace_bytes(size_t nq, size_t k, size_t ivf_nprobe);

    // ---- IVF prefilter (src/approx.rs through src/gpu.rs:435-462) ----
    // 0 keeps the reference's rule (choose_nlist, KMEANS_MAX_ITERS,
    // choose_nprobe); nlist and iters apply to the next build.
    void set_ivf_params(size_t nlist, size_t nprobe, size_t iters);
    // The IVF index, built now if needed (any index).
    const IvfIndex &build_ivf();
    // Install a saved IVF build instead of building one (checks the arrays;
    // fails if an IVF index already exists).  assign may be null.
    const IvfIndex &load_ivf(size_t nlist, const float *centroids, const uint64_t *offsets,
                             const uint64_t *orig, const uint64_t *assign);
    // approx_index (src/gpu.rs:435-441): the IVF index for approximate
    // searches, built on first use, or nullptr when the index was created
    // without approx or IVF would not speed it up.
    const IvfIndex *approx_index();
    // Enqueue an IVF search of device-resident queries (builds first if
    // needed); lens[q] = results found.
    void search_ivf_device(Workspace &ws, const float *dq, size_t nq, size_t k, size_t nprobe,
                           uint64_t *d_idx, float *d_sim, uint64_t *d_len, hipStream_t s);
    // Scan arguments of one IVF search over the workspace's buffers.
    IvfScanArgs ivf_scan_args(const IvfIndex &ivf, Workspace &ws, const float *dq, size_t nq,
                              size_t k, size_t nprobe, uint32_t kk, uint32_t qb, uint32_t nchunk,
                              uint64_t *d_idx, float *d_sim, uint64_t *d_len) const;
    // Queries per group of the IVF large-k form (their key arrays and their
    // selected keys within 256 MiB each) and the key array length per query.
    void ivf_large_k_shape(const IvfIndex &ivf, size_t nq, size_t nprobe, size_t k,
                           size_t &group, size_t &ld) const;
    // An index over rows already on the device (row-major, padded dims).
    static std::unique_ptr<GpuIndex> from_device_rows(const float *d_rows, size_t rows,
                                                      size_t padded_dims, int device);

    std::unique_ptr<Workspace> make_workspace(size_t max_queries, size_t max_k) const;
    uint32_t scan_grid() const;
    // Tiles of 64 rows each wave of the fused scan owns (0 = default).
    void set_scan_tiles_per_wave(uint32_t m) { scan_tpw_ = m; }
    void set_scan_flags(uint32_t f) { scan_flags_ = f; }
    uint32_t scan_tiles_per_wave() const;
    // Time the fused scan over candidate launch shapes and keep the fastest;
    // returns a one-line report per candidate.
    std::string autotune(int reps);
    // Times the batched MFMA scan's tile-shape variants per precision on
    // this index (its first rows as a 2048-query batch) and keeps the
    // fastest; results do not depend on the variant.
    std::string autotune_batch(int reps);
    void set_batch_variant(Precision p, int v) { mfma_variant_[static_cast<int>(p)] = v; }
    int batch_variant(Precision p) const { return mfma_variant_[static_cast<int>(p)]; }

  private:
    GpuIndex() = default;
    void init_device(int64_t device_index);
    void upload_tiled(const float *host, size_t rows, size_t src_dims, bool check_finite,
                      bool normalize, size_t logical_dims);
    uint64_t query_hash(const float *values, size_t n, size_t k, size_t query_rows,
                        bool approx) const;
    // raw_dims == 0: queries are normalised and padded already; else they
    // are prepared on the device in `mode` (kQuery*).
    std::vector<std::vector<Neighbor>> run_host(const float *queries, size_t query_rows,
                                                size_t k, size_t raw_dims, int mode, bool approx);

    int device_ = 0;
    uint32_t ncu_ = 0;
    uint32_t scan_tpw_ = 0;
    std::atomic<int> mfma_variant_[3] = {{0}, {0}, {0}};  // per Precision
    uint32_t scan_flags_ = kScanNonTemporal;
    size_t rows_ = 0, pdims_ = 0, d4_ = 0;
    bool approx_ = false, enable_cache_ = true;
    float4 *x_ = nullptr;
    uint4 *xbf16_ = nullptr;
    unsigned long long *xmax_sq_ = nullptr;
    unsigned long long *xerr_sq_ = nullptr;  // [0] unused, [1] BF16 rounding error
    std::mutex lazy_mu_;
    std::deque<RowSample> samples_;  // references stay valid as it grows
    const RowSample &sample_for(uint32_t stride, Precision p, hipStream_t s);
    void ensure_bf16(hipStream_t s);
    void ensure_fp8(hipStream_t s);
    void ensure_rowmajor(hipStream_t s);
    float4 *xrow_ = nullptr;                 // row-major FP32 copy (the reranks' reads)
    uint4 *xfp8_ = nullptr;                  // FP8-E4M3 copy, tiled units of 16
    float *xscale_ = nullptr;                // its per-row scales [rows_padded]
    unsigned long long *xerr8_sq_ = nullptr; // largest row rounding-error norm^2
    hipStream_t stream_ = nullptr;
    nndex_gpu_device_info info_{};
    std::mutex mu_;
    std::unique_ptr<Workspace> ws_;
    std::mutex ivf_mu_;
    std::unique_ptr<IvfIndex> ivf_;
    size_t ivf_nlist_ = 0, ivf_nprobe_ = 0, ivf_iters_ = 0;
    DeviceBuffer d_q_, d_raw_, d_idx_, d_sim_, d_flags_, d_len_;
    PinnedBuffer h_q_, h_idx_, h_sim_, h_flag_;
    std::unordered_map<uint64_t, std::vector<Neighbor>> cache_single_;
    std::unordered_map<uint64_t, std::vector<std::vector<Neighbor>>> cache_batch_;
};

void fill_device_info(int device, nndex_gpu_device_info &out);
int device_count();

// Thrown for front-end validation failures raised on the device.
void throw_flags(uint32_t flags);

}  // namespace nndex

// Opaque C handle of the backend boundary; shared with a front-end index.
struct nndex_gpu_index {
    std::shared_ptr<nndex::GpuIndex> impl;
};
