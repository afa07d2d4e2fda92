/*
 * nndex_hip.h -- C ABI of the MI355X (gfx950) search backend for nndex.
 *
 * Two layers are exported by libnndex_hip.so:
 *
 *  1. The backend boundary.  It replaces the crate-private `GpuIndex` of
 *     src/gpu.rs (wgpu + WGSL) one for one: the front-end hands over already
 *     normalised, zero-padded rows and queries, exactly as src/lib.rs:792-811,
 *     :353 and :399 do, and gets back owned (index, similarity) lists.  A Rust
 *     front-end binds these through `extern "C"`.
 *
 *  2. The front-end `NNdex` (src/lib.rs:160-599) restated in C++: validation,
 *     normalisation (on the device), k capping and the error model.  The
 *     Python package `nndex_amd` binds this layer.
 *
 * Plus a device-resident search entry point (inputs and outputs already in
 * HBM, work enqueued on a caller stream) and small device-memory / stream /
 * event helpers so callers need no other HIP binding.
 *
 * Conventions: every function returns an nndex_status; on failure a message
 * is kept per thread and returned by nndex_last_error().  Pointers are plain
 * host pointers unless the name says `device`.  Indices are uint64 (usize in
 * the reference), similarities float32.
 */
#ifndef NNDEX_HIP_H
#define NNDEX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  1..7 mirror NNdexError (src/lib.rs:107-145); 10..13 mirror
 * GpuError (src/gpu.rs:35-53), which the reference wraps in NNdexError::Gpu. */
typedef enum nndex_status {
    NNDEX_OK = 0,
    NNDEX_ERR_SHAPE_MISMATCH = 1,              /* NNdexError::ShapeMismatch */
    NNDEX_ERR_QUERY_DIM_MISMATCH = 2,          /* NNdexError::QueryDimensionalityMismatch */
    NNDEX_ERR_INVALID_TOP_K = 3,               /* NNdexError::InvalidTopK */
    NNDEX_ERR_NON_FINITE_INPUT = 4,            /* NNdexError::NonFiniteInput */
    NNDEX_ERR_ZERO_NORM = 5,                   /* NNdexError::ZeroNorm */
    NNDEX_ERR_CPU_BACKEND_UNAVAILABLE = 6,     /* NNdexError::CpuBackendUnavailable */
    NNDEX_ERR_GPU_BACKEND_UNAVAILABLE = 7,     /* NNdexError::GpuBackendUnavailable */
    NNDEX_ERR_GPU_ADAPTER_NOT_FOUND = 10,      /* GpuError::AdapterNotFound */
    NNDEX_ERR_GPU_REQUEST_DEVICE = 11,         /* GpuError::RequestDevice(String) */
    NNDEX_ERR_GPU_MAP_READ_FAILED = 12,        /* GpuError::MapReadFailed */
    NNDEX_ERR_GPU_DEVICE_INDEX_OUT_OF_RANGE = 13, /* GpuError::DeviceIndexOutOfRange */
    /* Boundary-only conditions (no reference counterpart). */
    NNDEX_ERR_INVALID_ARGUMENT = 20,
    NNDEX_ERR_OUT_OF_MEMORY = 21,
    NNDEX_ERR_HIP = 22,
    NNDEX_ERR_UNSUPPORTED = 23,
} nndex_status;

/* BackendPreference (src/lib.rs:46-51) / ActiveBackend (src/lib.rs:55-60). */
typedef enum nndex_backend { NNDEX_BACKEND_CPU = 0, NNDEX_BACKEND_GPU = 1 } nndex_backend;

/* IndexOptions (src/lib.rs:63-89).  gpu_device_index < 0 means None. */
typedef struct nndex_index_options {
    int32_t normalized;
    int32_t approx;
    int32_t backend; /* nndex_backend */
    int32_t enable_cache;
    int64_t gpu_device_index;
} nndex_index_options;

/* GpuDeviceInfo (src/gpu.rs:974-988); strings are NUL-terminated. */
typedef struct nndex_gpu_device_info {
    uint64_t index;
    char name[128];
    char driver[64];
    char driver_info[128];
    char backend[32];
    char device_type[32];
} nndex_gpu_device_info;

/* Message of the last failed call on this thread ("" if none). */
const char *nndex_last_error(void);
/* Library version string. */
const char *nndex_version(void);

/* ======================================================================== */
/* 1. Backend boundary (replaces GpuIndex, src/gpu.rs)                      */
/* ======================================================================== */

typedef struct nndex_gpu_index nndex_gpu_index;

/* GpuIndex::new (src/gpu.rs:138-145).  `storage` is rows x padded_dims f32,
 * row-major, already normalised and zero-padded; padded_dims % 8 == 0.
 * device_index < 0 selects the first device (automatic selection).
 * Errors: DEVICE_INDEX_OUT_OF_RANGE, ADAPTER_NOT_FOUND, OUT_OF_MEMORY. */
nndex_status nndex_gpu_index_new(const float *storage, size_t rows, size_t padded_dims,
                                 int32_t approx, int32_t enable_cache, int64_t device_index,
                                 nndex_gpu_index **out);
void nndex_gpu_index_free(nndex_gpu_index *index);

/* GpuIndex::search (src/gpu.rs:321-349).  `query` has padded_dims values,
 * normalised; `k` already capped by the caller (src/lib.rs:347).  Writes
 * min(k, rows) sorted (similarity desc, index asc) neighbours. */
nndex_status nndex_gpu_index_search(nndex_gpu_index *index, const float *query, size_t k,
                                    int32_t approx, uint64_t *out_indices,
                                    float *out_similarities, size_t *out_len);

/* GpuIndex::search_batch (src/gpu.rs:372-399).  Row q of the output starts at
 * q*k; out_lens[q] entries are valid. */
nndex_status nndex_gpu_index_search_batch(nndex_gpu_index *index, const float *queries,
                                          size_t query_rows, size_t k, int32_t approx,
                                          uint64_t *out_indices, float *out_similarities,
                                          size_t *out_lens);

/* GpuIndex::adapter_info (src/gpu.rs:312), reported as in NNdex::gpu_info
 * (src/lib.rs:302-318, index 0). */
nndex_status nndex_gpu_index_adapter_info(const nndex_gpu_index *index,
                                          nndex_gpu_device_info *out);

/* list_gpu_devices (src/gpu.rs:992-1010).  Writes up to `capacity` entries and
 * the total count; never fails for "no device" (count = 0). */
nndex_status nndex_list_gpu_devices(nndex_gpu_device_info *out, size_t capacity, size_t *count);

/* Shape of a built index. */
size_t nndex_gpu_index_rows(const nndex_gpu_index *index);
size_t nndex_gpu_index_padded_dims(const nndex_gpu_index *index);

/* ======================================================================== */
/* 2. Device-resident search (no reference counterpart: the reference always */
/*    round-trips through host buffers).  Inputs and outputs live in device  */
/*    memory; work is enqueued on `stream` (a hipStream_t, NULL = default)   */
/*    and the call returns without synchronising.  Calls on one index that   */
/*    overlap in time must use distinct workspaces.                          */
/* ======================================================================== */

typedef struct nndex_workspace nndex_workspace;

/* Scratch for up to `max_queries` queries of up to `max_k` neighbours. */
nndex_status nndex_workspace_new(const nndex_gpu_index *index, size_t max_queries, size_t max_k,
                                 nndex_workspace **out);
void nndex_workspace_free(nndex_workspace *ws);

/* Exact FP32 search of `query_rows` normalised padded queries at
 * `device_queries`; writes query_rows x k (index, similarity) to the device
 * outputs (k <= rows, k <= the workspace's max_k).  Up to 4 queries run the
 * fused scan without synchronising; larger batches run the certified MFMA
 * path (BF16 scoring up to 1024 padded dims, else FP32) and synchronise
 * `stream` once to re-run any uncertified query with the fused scan, so every
 * result is exact. */
nndex_status nndex_gpu_index_search_device(nndex_gpu_index *index, nndex_workspace *ws,
                                           const float *device_queries, size_t query_rows,
                                           size_t k, uint64_t *device_indices,
                                           float *device_similarities, void *stream);

/* Exact search that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as nndex_gpu_index_search_device (bit-identical
 * outputs), without a workspace: every query runs the fused scan (up to 4 per
 * launch), whose scratch is a static device array of 8 MiB of block-list keys
 * plus 1 KiB of counters per GPU, held until the process ends.  Rejects with
 * NNDEX_ERR_UNSUPPORTED a k whose block lists need more than that (scan grid
 * x keys per query above 1,048,576) and a k the fused scan cannot hold
 * (above 992; above 1984 when k equals rows), which the workspace entry point
 * serves with its large-k selection.  Counters are cleared by a
 * kernel at the start of each call.  Calls that overlap in time on one GPU
 * share the array and race: serialise them (one stream, or events).  Never
 * synchronises; no hipMalloc, hipFree or allocating library call. */
nndex_status nndex_gpu_index_search_device_noalloc(nndex_gpu_index *index,
                                                   const float *device_queries, size_t query_rows,
                                                   size_t k, uint64_t *device_indices,
                                                   float *device_similarities, void *stream);

/* Candidate-scoring precision of the batched path.  FP8 is OCP E4M3 with a
 * per-row scale max|x_r| / 448 (and a per-query scale likewise). */
typedef enum nndex_precision {
    NNDEX_PRECISION_F32 = 0,
    NNDEX_PRECISION_BF16 = 1,
    NNDEX_PRECISION_FP8 = 2
} nndex_precision;

/* Batched search: MFMA candidate scoring in `precision` (FP32 accumulation),
 * exact FP32 rerank of the best `overfetch` candidates (0 = default: k + 16
 * or 1.5k for F32, max(4k, k + 64) for BF16 and FP8; at most 447), and a
 * per-query certificate written to
 * device_flags (0 = result proven identical to the exact FP32 search, 1 =
 * not proven: re-run that query with nndex_gpu_index_search_device).  BF16
 * and F32 rescore every candidate above the threshold for queries the first
 * rerank does not prove; FP8 returns the exact rerank of the over-fetched
 * candidates as is (its certificate is reported, rarely reached).
 * Fully asynchronous on `stream`.  No reference counterpart (the reference
 * scores batches in FP32 only and reads every score back, src/gpu.rs:640-785). */
nndex_status nndex_gpu_index_search_batch_device(nndex_gpu_index *index, nndex_workspace *ws,
                                                 const float *device_queries, size_t query_rows,
                                                 size_t k, int32_t precision, size_t overfetch,
                                                 uint64_t *device_indices,
                                                 float *device_similarities,
                                                 uint8_t *device_flags, void *stream);

/* HIP graph of one batched search (no reference counterpart).  Records the
 * launches nndex_gpu_index_search_batch_device makes for exactly these
 * arguments (buffers, query_rows, k, precision, overfetch) on `stream` (not
 * the null stream) into an instantiated graph; every lazy copy and workspace
 * buffer is created first, so a replay allocates nothing.  Replay with
 * nndex_graph_launch after writing new queries into device_queries: the
 * outputs are those of a fresh batched search.  The workspace belongs to the
 * graph while it exists (no other call may use it concurrently). */
typedef struct nndex_graph nndex_graph;
nndex_status nndex_gpu_index_capture_batch(nndex_gpu_index *index, nndex_workspace *ws,
                                           const float *device_queries, size_t query_rows,
                                           size_t k, int32_t precision, size_t overfetch,
                                           uint64_t *device_indices, float *device_similarities,
                                           uint8_t *device_flags, void *stream,
                                           nndex_graph **out);
nndex_status nndex_graph_launch(nndex_graph *graph, void *stream);
void nndex_graph_free(nndex_graph *graph);

/* Autotuner (no reference counterpart): times the fused scan over candidate
 * launch shapes (64-row tiles per wavefront) on this index and keeps the
 * fastest; writes a text report.  Workspaces made before the call must be
 * re-created. */
nndex_status nndex_gpu_index_autotune(nndex_gpu_index *index, int32_t reps, char *report,
                                      size_t report_len);
/* Batch autotuner (no reference counterpart): times the batched MFMA scan's
 * tile-shape variants (rows per wavefront, queries per LDS chunk) for each
 * scoring precision the index supports, on a 2048-query batch of its own rows,
 * and keeps the fastest per precision; writes a text report.  Results do not
 * depend on the variant.  HIP graphs captured before keep their shapes. */
nndex_status nndex_gpu_index_autotune_batch(nndex_gpu_index *index, int32_t reps, char *report,
                                            size_t report_len);
/* Force a batched scan tile-shape variant (0 = default, 1 = the alternative;
 * a row width without an alternative runs the default) for one precision
 * (NNDEX_PRECISION_*). */
nndex_status nndex_gpu_index_set_batch_tiles(nndex_gpu_index *index, int32_t precision,
                                             int32_t variant);
/* Set the fused scan's launch shape: 64-row tiles per wavefront (0 =
 * default) and flags (bit 0: non-temporal loads, bit 1: contiguous tile runs
 * per wavefront instead of interleaved tiles, bits 2-3: single-query load
 * group of 8 / 12 / 16 float4 per lane, two groups in flight). */
nndex_status nndex_gpu_index_set_launch(nndex_gpu_index *index, uint32_t scan_tiles_per_wave,
                                        uint32_t scan_flags);

/* ---- IVF prefilter (src/approx.rs, used by src/gpu.rs:435-462) ---------- */

/* Built IVF index summary.  `rechecked` counts the assignment rows (summed
 * over iterations) whose top-two MFMA scores were too close to separate and
 * were re-evaluated in the reference dot order. */
typedef struct nndex_ivf_info {
    uint64_t rows;
    uint64_t padded_dims;
    uint64_t nlist;
    uint64_t iters;
    uint64_t rechecked;
    double build_ms;
} nndex_ivf_info;

/* IVF parameters (replaces the constants of src/approx.rs: choose_nlist
 * :305-311, KMEANS_MAX_ITERS :28, choose_nprobe :318-336); 0 keeps the
 * reference's rule for that field.  nlist and iters apply to the next build
 * only, so they must be set before the first approximate search. */
nndex_status nndex_gpu_index_set_ivf_params(nndex_gpu_index *index, size_t nlist, size_t nprobe,
                                            size_t iters);
/* Build the IVF index now on the device (ApproxIndex::build,
 * src/approx.rs:46-100; the reference builds lazily on the first approximate
 * search, src/gpu.rs:435-441, which nndex_gpu_index_search also does).  A
 * second call returns the existing build.  Works for any index; searches
 * with approx = 1 use it only if the index was created with approx = 1. */
nndex_status nndex_gpu_index_build_ivf(nndex_gpu_index *index, nndex_ivf_info *out);
/* Copy the built IVF index to host memory; any pointer may be NULL.
 * centroids: nlist x padded_dims; offsets: nlist + 1 (cluster_offsets);
 * original_indices: rows (reordered position -> row); assignments: rows. */
nndex_status nndex_gpu_index_ivf_export(nndex_gpu_index *index, float *centroids,
                                        uint64_t *offsets, uint64_t *original_indices,
                                        uint64_t *assignments);
/* Install a saved IVF build (the arrays nndex_gpu_index_ivf_export writes,
 * same layouts; assignments may be NULL) instead of building one, so an index
 * re-created over the same rows skips the k-means build.  The reference keeps
 * no saved form (ApproxIndex, src/approx.rs:11-25, is rebuilt lazily on each
 * construction, src/gpu.rs:435-441); this is the persistence that build
 * lacks.  Checks every array (offsets from 0 to rows and non-decreasing, a
 * permutation, assignments consistent with both, finite centroids) and fails
 * with NNDEX_ERR_INVALID_ARGUMENT if the index already has an IVF index.
 * The reordered copy is gathered again from the resident rows; `out` (may be
 * NULL) reports iters = 0 and build_ms = the import time. */
nndex_status nndex_gpu_index_ivf_import(nndex_gpu_index *index, size_t nlist,
                                        const float *centroids, const uint64_t *offsets,
                                        const uint64_t *original_indices,
                                        const uint64_t *assignments, nndex_ivf_info *out);
/* Device-resident approximate search (candidate_indices + rerank,
 * src/approx.rs:110-138 and src/gpu.rs:448-462): the nprobe (0 = the
 * reference's choose_nprobe for k) clusters whose centroids score highest,
 * then every row of those clusters rescored exactly into a top-k.  Writes
 * query_rows x k results and device_lens[q] = results found (< k when the
 * probed clusters hold fewer than k rows).  Builds the IVF index first if
 * needed.  Synchronises `stream` once when query_rows > 4. */
nndex_status nndex_gpu_index_search_ivf_device(nndex_gpu_index *index, nndex_workspace *ws,
                                               const float *device_queries, size_t query_rows,
                                               size_t k, size_t nprobe,
                                               uint64_t *device_indices,
                                               float *device_similarities,
                                               uint64_t *device_lens, void *stream);

/* ---- Serving loop (no reference counterpart: the reference serves one  */
/* call at a time behind a Mutex, src/gpu.rs:511-592) ---------------------- */

/* Continuous batching over one resident index.  Requests (normalised padded
 * query rows, k, mode 0 = exact / 1 = IVF with nprobe) are queued; each of
 * `slots` executors (own stream, workspace and staging, sized once for
 * max_batch_queries x max_k) takes the oldest pending request plus every
 * later one of the same kind that fits and runs them as one batch at their
 * largest k, so work arriving while batches run is admitted as executors free
 * and overlaps on the device; with use_graphs, exact batches replay captured
 * HIP graphs (one per rows/k bucket and executor).  Admission by HBM
 * footprint: the number of
 * executors is the largest whose fixed footprint fits hbm_budget_bytes;
 * nothing is allocated while serving.  A request's results equal those of the
 * same request served alone. */
typedef struct nndex_server_options {
    uint32_t slots;             /* executors in flight; 0 = 4 */
    uint32_t max_batch_queries; /* rows per batch; 0 = 2048 (larger requests are split) */
    uint32_t max_k;             /* largest k served; 0 = 100 */
    uint32_t ivf_nprobe_max;    /* 0 = exact requests only; else IVF up to this nprobe */
    uint64_t hbm_budget_bytes;  /* device bytes for all executors; 0 = 16 GiB */
    uint32_t use_graphs;        /* 1 = exact batches above 4 rows replay HIP
                                   graphs captured per (rows, k) bucket: rows
                                   to a multiple of 16 (<= 128) or 64, k to a
                                   power of two; 0 = eager launches */
} nndex_server_options;

typedef struct nndex_server_stats {
    uint64_t slots;             /* executors admitted by the budget */
    uint64_t slot_bytes;        /* device footprint of one executor */
    uint64_t requests;          /* completed since the last reset */
    uint64_t queries;           /* rows of those requests */
    uint64_t batches;
    double mean_batch_queries;
    double p50_ms, p99_ms, max_ms;  /* submit -> results ready, per request */
    double span_s;              /* first submit -> last completion */
    double queries_per_s, requests_per_s;
    uint64_t graph_batches;     /* batches served by a captured graph */
} nndex_server_stats;

typedef struct nndex_server nndex_server;

nndex_status nndex_server_new(nndex_gpu_index *index, const nndex_server_options *options,
                              nndex_server **out);
/* Stops the executors after their current batch; pending requests are dropped. */
void nndex_server_free(nndex_server *server);
/* Queue `query_rows` normalised padded rows; returns a ticket at once.
 * nprobe = 0 picks the reference's choose_nprobe for k (IVF only). */
nndex_status nndex_server_submit(nndex_server *server, const float *queries, size_t query_rows,
                                 size_t k, int32_t mode, size_t nprobe, uint64_t *ticket);
/* Queue `query_rows` raw rows of `dims` values each (dims rounded up to a
 * multiple of 8 must equal the index's padded dims), as the front-end's
 * search_batch receives them (src/lib.rs:372-402): normalize = 1 normalises
 * each row on the device as normalize_rows does (src/lib.rs:461-553), 0 only
 * checks and pads (ensure_finite + pad_rows, for an index of normalized
 * rows).  The checks run on the device with the batch; a row that fails them
 * fails its request only, reported by nndex_server_wait as
 * NNDEX_ERR_NON_FINITE_INPUT or NNDEX_ERR_ZERO_NORM.  Raw requests batch with
 * raw requests of the same dims and normalize. */
nndex_status nndex_server_submit_raw(nndex_server *server, const float *queries,
                                     size_t query_rows, size_t dims, int32_t normalize, size_t k,
                                     int32_t mode, size_t nprobe, uint64_t *ticket);
/* Block until the ticket's request is done; copy query_rows x min(k, rows)
 * indices and similarities and query_rows result counts (any may be NULL). */
nndex_status nndex_server_wait(nndex_server *server, uint64_t ticket, uint64_t *out_indices,
                               float *out_similarities, uint64_t *out_lens,
                               double *latency_ms);
nndex_status nndex_server_get_stats(nndex_server *server, nndex_server_stats *out);
nndex_status nndex_server_reset_stats(nndex_server *server);

/* ======================================================================== */
/* 3. Front-end NNdex (src/lib.rs) in C++                                   */
/* ======================================================================== */

typedef struct nndex_index nndex_index;

/* NNdex::new (src/lib.rs:185-278): `matrix` holds `values` floats. */
nndex_status nndex_new(const float *matrix, size_t values, size_t rows, size_t dims,
                       const nndex_index_options *options, nndex_index **out);
void nndex_free(nndex_index *index);
/* NNdex::backend / rows / dims (src/lib.rs:281-299). */
int32_t nndex_backend_of(const nndex_index *index);
size_t nndex_rows(const nndex_index *index);
size_t nndex_dims(const nndex_index *index);
/* NNdex::gpu_info (src/lib.rs:302-318); *has_info = 0 for a CPU index. */
nndex_status nndex_gpu_info(const nndex_index *index, nndex_gpu_device_info *out,
                            int32_t *has_info);
/* NNdex::search (src/lib.rs:335-355): query has `query_len` values. */
nndex_status nndex_search(nndex_index *index, const float *query, size_t query_len, size_t k,
                          uint64_t *out_indices, float *out_similarities, size_t *out_len);
/* NNdex::search_batch (src/lib.rs:372-402): queries has `values` floats.
 * Output rows are strided by min(k, rows). */
nndex_status nndex_search_batch(nndex_index *index, const float *queries, size_t values,
                                size_t query_rows, size_t k, uint64_t *out_indices,
                                float *out_similarities, size_t *out_lens);
/* The backend handle behind a GPU index (NULL for CPU). */
nndex_gpu_index *nndex_backend_handle(nndex_index *index);

/* ======================================================================== */
/* 4. Device helpers                                                        */
/* ======================================================================== */

nndex_status nndex_device_count(int32_t *count);
nndex_status nndex_set_device(int32_t device);
nndex_status nndex_device_malloc(void **ptr, size_t bytes);
nndex_status nndex_device_free(void *ptr);
nndex_status nndex_memcpy_h2d(void *device_dst, const void *host_src, size_t bytes);
nndex_status nndex_memcpy_d2h(void *host_dst, const void *device_src, size_t bytes);
nndex_status nndex_stream_create(void **stream);
nndex_status nndex_stream_destroy(void *stream);
nndex_status nndex_stream_synchronize(void *stream);
nndex_status nndex_event_create(void **event);
nndex_status nndex_event_destroy(void *event);
nndex_status nndex_event_record(void *event, void *stream);
nndex_status nndex_event_synchronize(void *event);
nndex_status nndex_event_elapsed_ms(void *start, void *stop, float *ms);
/* Plain streaming copy kernel (bandwidth calibration). */
nndex_status nndex_device_copy(void *device_dst, const void *device_src, size_t bytes,
                               void *stream);
/* Diagnostic with no counterpart in the reference, for tests of the batched
 * path's certificate: one wavefront runs `calls` chained MFMAs of the
 * candidate scan's element form `elem` (0 FP32 16x16x4, 1 BF16 16x16x32,
 * 2 FP8-E4M3 16x16x128 with unit scales) over host operands laid out per
 * call and lane (FP32: one float; BF16: 8 values; FP8: 32 codes; lane l
 * holds row / column l % 16 and K block l / 16), and writes the 16 x 16 FP32
 * result row-major to host `out`.  Synchronous on the current device. */
nndex_status nndex_debug_mfma_tile(int32_t elem, const void *a, const void *b, uint32_t calls,
                                   float *out);

#ifdef __cplusplus
}
#endif

#endif /* NNDEX_HIP_H */
