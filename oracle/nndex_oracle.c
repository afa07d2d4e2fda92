/*
 * nndex_oracle.c -- TEST INFRASTRUCTURE ONLY: CPU restatement of nndex's exact
 * search path (fblissjr/nndex v0.2.1) plus an FP64 brute-force oracle.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as
 * the checker or as the timed CPU baseline; the product never links it.
 *
 * What pins each part (see tests/test_oracle_*.py):
 *   - orc_dot_product: restates simsimd ^6.5.12 `simsimd_dot_f32_neon` (four f32
 *     FMA lanes over 4-element strides, pairwise lane reduction, scalar tail).
 *     PINNED bit-for-bit by the published golden vector of the 50,000x128 k=5
 *     case (notebooks/python_bindings_demo.ipynb:102-103): all five scores
 *     reproduce exactly; the f64-widening x86 variant misses one by 1 ulp.
 *   - orc_normalize_rows / orc_normalize_vector: lib.rs:461-575, pinned by the
 *     same golden vector and by the closed-form KATs of lib.rs:1335-1352.
 *   - TopK accumulator and topk_from_scores: topk.rs:44-210, pinned by the 20
 *     KATs of topk.rs:216-477 restated in tests, and by brute-force sorting.
 *   - strategy selection: profiles.rs:72-185, pinned by profiles.rs:194-343.
 *   - orc_gpu_exact_search: gpu.rs:575-589 / :758-776 (one accumulator, rows
 *     pushed in order); its choice among rows tied at the k-th score pinned by
 *     a closed-form restatement of the heap and by a hand-worked case.
 *   - orc_sgemm_scores: restates matrixmultiply's f32 kernel (sequential-k FMA
 *     within KC=256 blocks, block partials added into C).  Bitwise parity
 *     UNPINNED (matrixmultiply is not vendored; the reference's macOS build may
 *     use Accelerate); pinned only to 1e-5 against FP64 and to batch==single.
 *   - orc_dot_ndarray: ndarray's unrolled_dot (eight f32 accumulators), used by
 *     approx.rs GEMVs.  Bitwise parity UNPINNED; pinned to 1e-5 against FP64.
 *   - rayon reduce trees (ParallelChunked/ParallelFold/MatrixChunked) are
 *     restated as a left-to-right merge of contiguous segments.  Results differ
 *     from any particular rayon run only when scores tie exactly.
 *   - IVF build (approx.rs:46-552): choose_nlist/nprobe/isqrt pinned by
 *     approx.rs:567-599; layout invariants by approx.rs:618-644; recall floors
 *     by lib.rs:961-1021, 1204-1297.
 *   - FP64 oracle: textbook f64 normalise + dot + full sort; pinned by
 *     closed forms (orthonormal bases, [3,4] -> unit) in tests.
 */
#include "nndex_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static int resolve_threads(int threads) {
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_num_procs();
#else
    threads = 1;
#endif
    return threads < 1 ? 1 : threads;
}

/* ------------------------------------------------------------------------- */
/* f32 total order (Rust f32::total_cmp) and heap items (topk.rs:13-37)       */
/* ------------------------------------------------------------------------- */

static inline int32_t total_key(float f) {
    int32_t bits;
    memcpy(&bits, &f, 4);
    bits ^= (int32_t)(((uint32_t)(bits >> 31)) >> 1);
    return bits;
}

typedef struct {
    uint64_t index;
    float similarity;
} item_t;

/* HeapItem::cmp: similarity.total_cmp, then index (topk.rs:25-31). */
static inline int item_cmp(item_t a, item_t b) {
    int32_t ka = total_key(a.similarity), kb = total_key(b.similarity);
    if (ka != kb) return ka < kb ? -1 : 1;
    if (a.index != b.index) return a.index < b.index ? -1 : 1;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Validation, normalisation, padding (lib.rs:433-599)                       */
/* ------------------------------------------------------------------------- */

/* ensure_finite, lib.rs:445-450 */
int orc_ensure_finite(const float *values, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!isfinite(values[i])) return ORC_NON_FINITE;
    return ORC_OK;
}

/* pad_rows, lib.rs:578-588 */
void orc_pad_rows(const float *values, size_t rows, size_t dims, size_t padded_dims, float *out) {
    for (size_t r = 0; r < rows; ++r) {
        memcpy(out + r * padded_dims, values + r * dims, dims * sizeof(float));
        for (size_t c = dims; c < padded_dims; ++c) out[r * padded_dims + c] = 0.0f;
    }
}

/* simsimd_dot_f32_neon restated: four FMA lanes, vaddvq_f32 pairwise reduce,
 * scalar tail (clang contracts `ab += a*b` into an FMA on that target). */
float orc_dot_product_simd(const float *a, const float *b, size_t n) {
    float l0 = 0.f, l1 = 0.f, l2 = 0.f, l3 = 0.f;
    size_t i = 0;
    for (; i + 4 <= n; i += 4) {
        l0 = fmaf(a[i + 0], b[i + 0], l0);
        l1 = fmaf(a[i + 1], b[i + 1], l1);
        l2 = fmaf(a[i + 2], b[i + 2], l2);
        l3 = fmaf(a[i + 3], b[i + 3], l3);
    }
    float ab = (l0 + l1) + (l2 + l3);
    for (; i < n; ++i) ab = fmaf(a[i], b[i], ab);
    return ab;
}

/* dot_unrolled_8, cpu.rs:667-695 (no FMA contraction: Rust never contracts). */
float orc_dot_unrolled_8(const float *lhs, const float *rhs, size_t n) {
    float s[8] = {0};
    size_t idx = 0;
    while (idx + 8 <= n) {
        for (int l = 0; l < 8; ++l) s[l] += lhs[idx + l] * rhs[idx + l];
        idx += 8;
    }
    while (idx < n) {
        s[0] += lhs[idx] * rhs[idx];
        idx += 1;
    }
    return ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}

/* dot_product, cpu.rs:653-660: simsimd when len >= 8, else dot_unrolled_8. */
float orc_dot_product(const float *lhs, const float *rhs, size_t n) {
    if (n >= 8) return orc_dot_product_simd(lhs, rhs, n);
    return orc_dot_unrolled_8(lhs, rhs, n);
}

/* ndarray numeric_util::unrolled_dot (eight accumulators, pairwise (p_i+p_{i+4})). */
float orc_dot_ndarray(const float *xs, const float *ys, size_t n) {
    float p[8] = {0};
    size_t i = 0;
    for (; i + 8 <= n; i += 8)
        for (int l = 0; l < 8; ++l) p[l] = p[l] + xs[i + l] * ys[i + l];
    float sum = 0.f;
    sum = sum + (p[0] + p[4]);
    sum = sum + (p[1] + p[5]);
    sum = sum + (p[2] + p[6]);
    sum = sum + (p[3] + p[7]);
    for (; i < n; ++i) sum = sum + xs[i] * ys[i];
    return sum;
}

/* dot_small, approx.rs:538-551 */
static float dot_small(const float *lhs, const float *rhs, size_t n) {
    if (n >= 8) return orc_dot_product_simd(lhs, rhs, n);
    float sum = 0.f;
    for (size_t i = 0; i < n; ++i) sum += lhs[i] * rhs[i];
    return sum;
}

/* normalize_row_into, lib.rs:519-553 */
static int normalize_row_into(const float *src, float *out_row, size_t dims, size_t padded_dims,
                              int prefer_simd_norm) {
    float norm_sq;
    if (prefer_simd_norm) {
        norm_sq = orc_dot_product_simd(src, src, dims);
        if (!isfinite(norm_sq)) return ORC_NON_FINITE;
    } else {
        norm_sq = 0.f;
        for (size_t i = 0; i < dims; ++i) {
            if (!isfinite(src[i])) return ORC_NON_FINITE;
            norm_sq += src[i] * src[i];
        }
    }
    if (norm_sq <= FLT_EPSILON) return ORC_ZERO_NORM;
    float inv_norm = 1.0f / sqrtf(norm_sq);
    for (size_t i = 0; i < dims; ++i) out_row[i] = src[i] * inv_norm;
    for (size_t i = dims; i < padded_dims; ++i) out_row[i] = 0.0f;
    return ORC_OK;
}

/* normalize_rows, lib.rs:461-514.  Rows are processed in order and the first
 * failing row's error is returned (rayon's try_for_each reports an arbitrary
 * failing row when several fail). */
int orc_normalize_rows(const float *values, size_t rows, size_t dims, size_t padded_dims,
                       float *out) {
    int prefer_simd_norm = dims >= 16;
    for (size_t r = 0; r < rows; ++r) {
        int st = normalize_row_into(values + r * dims, out + r * padded_dims, dims, padded_dims,
                                    prefer_simd_norm);
        if (st != ORC_OK) return st;
    }
    return ORC_OK;
}

/* normalize_vector, lib.rs:558-575: scalar sequential f32 sum of squares. */
int orc_normalize_vector(const float *values, size_t dims, size_t padded_dims, float *out) {
    float norm_sq = 0.f;
    for (size_t i = 0; i < dims; ++i) {
        if (!isfinite(values[i])) return ORC_NON_FINITE;
        norm_sq += values[i] * values[i];
    }
    if (norm_sq <= FLT_EPSILON) return ORC_ZERO_NORM;
    float inv_norm = 1.0f / sqrtf(norm_sq);
    for (size_t i = 0; i < dims; ++i) out[i] = values[i] * inv_norm;
    for (size_t i = dims; i < padded_dims; ++i) out[i] = 0.0f;
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* TopKAccumulator (topk.rs:44-152): std::collections::BinaryHeap semantics    */
/* over Reverse<HeapItem>, i.e. the root is the smallest HeapItem.            */
/* ------------------------------------------------------------------------- */

struct orc_topk {
    size_t k;
    size_t len;
    item_t *heap; /* capacity k */
    int has_best;
    item_t best;
    float min_threshold;
};

/* In the max-heap of Reverse<HeapItem>, x "<=" y  <=>  HeapItem x >= y. */
static inline int rev_le(item_t x, item_t y) { return item_cmp(x, y) >= 0; }
static inline int rev_lt(item_t x, item_t y) { return item_cmp(x, y) > 0; }
static inline int rev_ge(item_t x, item_t y) { return item_cmp(x, y) <= 0; }

/* BinaryHeap::sift_up(0, pos) */
static void heap_sift_up(item_t *data, size_t pos) {
    item_t elem = data[pos];
    while (pos > 0) {
        size_t parent = (pos - 1) / 2;
        if (rev_le(elem, data[parent])) break;
        data[pos] = data[parent];
        pos = parent;
    }
    data[pos] = elem;
}

/* BinaryHeap::sift_down_range(0, end), run when a PeekMut is dropped. */
static void heap_sift_down(item_t *data, size_t end) {
    size_t pos = 0;
    item_t elem = data[0];
    size_t child = 1;
    while (end >= 2 && child <= end - 2) {
        child += rev_le(data[child], data[child + 1]) ? 1 : 0;
        if (rev_ge(elem, data[child])) {
            data[pos] = elem;
            return;
        }
        data[pos] = data[child];
        pos = child;
        child = 2 * pos + 1;
    }
    if (end >= 1 && child == end - 1 && rev_lt(elem, data[child])) {
        data[pos] = data[child];
        pos = child;
    }
    data[pos] = elem;
}

orc_topk *orc_topk_new(size_t k) {
    orc_topk *acc = (orc_topk *)calloc(1, sizeof(orc_topk));
    acc->k = k;
    acc->heap = (item_t *)malloc((k ? k : 1) * sizeof(item_t));
    acc->min_threshold = -INFINITY;
    return acc;
}

void orc_topk_free(orc_topk *acc) {
    if (!acc) return;
    free(acc->heap);
    free(acc);
}

float orc_topk_min_threshold(const orc_topk *acc) { return acc->min_threshold; }
size_t orc_topk_len(const orc_topk *acc) { return acc->k == 1 ? (size_t)acc->has_best : acc->len; }

/* push_slow, topk.rs:80-105 */
void orc_topk_push_slow(orc_topk *acc, uint64_t index, float similarity) {
    item_t it = {index, similarity};
    if (acc->k == 1) {
        acc->best = it;
        acc->has_best = 1;
        acc->min_threshold = similarity;
        return;
    }
    if (acc->k == 0) return;
    if (acc->len >= acc->k) {
        acc->heap[0] = it;
        heap_sift_down(acc->heap, acc->len);
        acc->min_threshold = acc->heap[0].similarity;
        return;
    }
    acc->heap[acc->len] = it;
    heap_sift_up(acc->heap, acc->len);
    acc->len += 1;
    if (acc->len == acc->k) acc->min_threshold = acc->heap[0].similarity;
}

/* push, topk.rs:70-75: `<=` fast reject */
void orc_topk_push(orc_topk *acc, uint64_t index, float similarity) {
    if (similarity <= acc->min_threshold) return;
    orc_topk_push_slow(acc, index, similarity);
}

/* merge, topk.rs:108-119: pushes other's heap in its storage order. */
void orc_topk_merge(orc_topk *acc, orc_topk *other) {
    if (acc->k == 1) {
        if (other->has_best) orc_topk_push(acc, other->best.index, other->best.similarity);
    } else {
        for (size_t i = 0; i < other->len; ++i)
            orc_topk_push(acc, other->heap[i].index, other->heap[i].similarity);
    }
    orc_topk_free(other);
}

/* Sort by (similarity desc via total_cmp, index asc): topk.rs:143-148. */
static int out_cmp(const void *pa, const void *pb) {
    const item_t *a = (const item_t *)pa, *b = (const item_t *)pb;
    int32_t ka = total_key(a->similarity), kb = total_key(b->similarity);
    if (ka != kb) return ka > kb ? -1 : 1;
    if (a->index != b->index) return a->index < b->index ? -1 : 1;
    return 0;
}

/* into_sorted_vec, topk.rs:122-151 (consumes). */
size_t orc_topk_into_sorted(orc_topk *acc, uint64_t *idx, float *sim) {
    size_t n = 0;
    if (acc->k == 1) {
        if (acc->has_best) {
            idx[0] = acc->best.index;
            sim[0] = acc->best.similarity;
            n = 1;
        }
    } else {
        qsort(acc->heap, acc->len, sizeof(item_t), out_cmp);
        for (size_t i = 0; i < acc->len; ++i) {
            idx[i] = acc->heap[i].index;
            sim[i] = acc->heap[i].similarity;
        }
        n = acc->len;
    }
    orc_topk_free(acc);
    return n;
}

/* topk_from_scores, topk.rs:169-210 */
size_t orc_topk_from_scores(const float *scores, size_t n, size_t k, size_t row_offset,
                            uint64_t *idx, float *sim) {
    size_t capped_k = k < n ? k : n;
    if (capped_k == 0) return 0;
    if (n > 131072) {
        orc_topk *acc = orc_topk_new(capped_k);
        float threshold = acc->min_threshold;
        for (size_t i = 0; i < n; ++i) {
            if (scores[i] > threshold) {
                orc_topk_push_slow(acc, i + row_offset, scores[i]);
                threshold = acc->min_threshold;
            }
        }
        return orc_topk_into_sorted(acc, idx, sim);
    }
    /* select_nth_unstable_by + sort under the total order == full sort, take k */
    item_t *items = (item_t *)malloc(n * sizeof(item_t));
    for (size_t i = 0; i < n; ++i) {
        items[i].index = i;
        items[i].similarity = scores[i];
    }
    qsort(items, n, sizeof(item_t), out_cmp);
    for (size_t i = 0; i < capped_k; ++i) {
        idx[i] = items[i].index + row_offset;
        sim[i] = items[i].similarity;
    }
    free(items);
    return capped_k;
}

/* ------------------------------------------------------------------------- */
/* Strategy selection (profiles.rs:72-185, cpu.rs:697-721)                   */
/* ------------------------------------------------------------------------- */

size_t orc_profile_cutoff(size_t rows, size_t dims, size_t query_rows) {
    unsigned __int128 w = (unsigned __int128)rows * dims * (query_rows ? query_rows : 1);
    if (w >= 200000000u) return 1024;
    if (w >= 20000000u) return 2048;
    return 8192;
}

int orc_cpu_search_strategy(size_t profile_cutoff, size_t rows, size_t dims,
                            size_t parallel_cutoff_rows, size_t chunk_rows,
                            size_t scores_vec_max_rows, int has_blas) {
    size_t eff = profile_cutoff < parallel_cutoff_rows ? profile_cutoff : parallel_cutoff_rows;
    if (rows <= eff) {
        if (has_blas && dims >= 64 && rows >= 128) return ORC_S_GEMV;
        return ORC_S_SERIAL;
    }
    if (rows <= scores_vec_max_rows) {
        if (has_blas && dims >= 64) return ORC_S_GEMV;
        return ORC_S_PARALLEL_SCORES;
    }
    if (chunk_rows >= 128) return ORC_S_PARALLEL_CHUNKED;
    return ORC_S_PARALLEL_FOLD;
}

int orc_cpu_batch_strategy(size_t rows, size_t dims, size_t query_rows, int prefer_query_parallel) {
    int gemm_eligible = (dims >= 64 && query_rows >= 4) || (dims >= 32 && query_rows >= 8) ||
                        (dims >= 16 && query_rows >= 16);
    if (gemm_eligible) {
        unsigned __int128 bytes = (unsigned __int128)query_rows * rows * 4u;
        if (bytes <= 128u * 1024u * 1024u) return ORC_B_GEMM;
        return ORC_B_GEMM_CHUNKED;
    }
    unsigned __int128 mbytes = (unsigned __int128)rows * dims * 4u;
    if (query_rows >= 4 && mbytes > 64u * 1024u * 1024u) return ORC_B_MATRIX_CHUNKED;
    if (prefer_query_parallel) return ORC_B_QUERY_PARALLEL;
    return ORC_B_PROFILED_PER_QUERY;
}

size_t orc_parallel_cutoff(size_t dims, int threads) {
    size_t per = 32768 / (dims ? dims : 1);
    if (per < 1) per = 1;
    return (size_t)resolve_threads(threads) * per;
}

size_t orc_chunk_rows(size_t rows, size_t dims) {
    size_t target = 32768 > dims ? 32768 : dims;
    size_t r = target / dims;
    if (r < 1) r = 1;
    size_t cap = rows ? rows : 1;
    return r < cap ? r : cap;
}

size_t orc_batch_chunk_rows(size_t dims) {
    size_t r = (256 * 1024) / (dims * 4);
    return r > 64 ? r : 64;
}

/* ------------------------------------------------------------------------- */
/* Exact CPU strategies (cpu.rs:216-639)                                     */
/* ------------------------------------------------------------------------- */

#define SCORES_VEC_MAX_ROWS 4096

typedef struct {
    const float *matrix;
    size_t rows, dims;
    int threads;
} cpu_ctx;

/* search_serial, cpu.rs:555-562 */
static orc_topk *search_serial(const cpu_ctx *c, const float *q, size_t k) {
    orc_topk *acc = orc_topk_new(k);
    for (size_t r = 0; r < c->rows; ++r)
        orc_topk_push(acc, r, orc_dot_product(c->matrix + r * c->dims, q, c->dims));
    return acc;
}

/* search_parallel_scores, cpu.rs:529-552 */
static orc_topk *search_parallel_scores(const cpu_ctx *c, const float *q, size_t k) {
    float *scores = (float *)malloc(c->rows * sizeof(float));
#pragma omp parallel for num_threads(c->threads) schedule(static)
    for (size_t r = 0; r < c->rows; ++r)
        scores[r] = orc_dot_product(c->matrix + r * c->dims, q, c->dims);
    orc_topk *acc = orc_topk_new(k);
    for (size_t r = 0; r < c->rows; ++r) orc_topk_push(acc, r, scores[r]);
    free(scores);
    return acc;
}

/* Shared by search_parallel_chunked (cpu.rs:594-617) and search_parallel_fold
 * (cpu.rs:620-639): contiguous segments, each folded with push, merged left to
 * right. */
static orc_topk *search_segments(const cpu_ctx *c, const float *q, size_t k, size_t seg_rows) {
    size_t nseg = (c->rows + seg_rows - 1) / seg_rows;
    orc_topk **parts = (orc_topk **)malloc(nseg * sizeof(orc_topk *));
#pragma omp parallel for num_threads(c->threads) schedule(dynamic, 4)
    for (size_t s = 0; s < nseg; ++s) {
        size_t r0 = s * seg_rows, r1 = r0 + seg_rows < c->rows ? r0 + seg_rows : c->rows;
        orc_topk *acc = orc_topk_new(k);
        for (size_t r = r0; r < r1; ++r)
            orc_topk_push(acc, r, orc_dot_product(c->matrix + r * c->dims, q, c->dims));
        parts[s] = acc;
    }
    orc_topk *acc = orc_topk_new(k);
    for (size_t s = 0; s < nseg; ++s) orc_topk_merge(acc, parts[s]);
    free(parts);
    return acc;
}

static orc_topk *search_with_profile(const cpu_ctx *c, const float *q, size_t k,
                                     size_t profile_cutoff, int *strategy_out) {
    size_t chunked = orc_chunk_rows(c->rows, c->dims);
    int s = orc_cpu_search_strategy(profile_cutoff, c->rows, c->dims,
                                    orc_parallel_cutoff(c->dims, c->threads), chunked,
                                    SCORES_VEC_MAX_ROWS, 0);
    if (strategy_out) *strategy_out = s;
    switch (s) {
    case ORC_S_SERIAL:
        return search_serial(c, q, k);
    case ORC_S_PARALLEL_SCORES:
        return search_parallel_scores(c, q, k);
    case ORC_S_PARALLEL_CHUNKED:
        return search_segments(c, q, k, chunked);
    default: {
        /* ParallelFold: rayon splits rows adaptively; restated as 8 segments
         * per thread. */
        size_t nseg = (size_t)c->threads * 8;
        size_t seg = (c->rows + nseg - 1) / nseg;
        if (seg < 1) seg = 1;
        return search_segments(c, q, k, seg);
    }
    }
}

size_t orc_cpu_search(const float *matrix, size_t rows, size_t dims, const float *query,
                      size_t k, int threads, uint64_t *idx, float *sim, int *strategy_out) {
    cpu_ctx c = {matrix, rows, dims, resolve_threads(threads)};
    size_t capped = k < rows ? k : rows;
    orc_topk *acc =
        search_with_profile(&c, query, capped, orc_profile_cutoff(rows, dims, 1), strategy_out);
    return orc_topk_into_sorted(acc, idx, sim);
}

/* GPU backend's exact search (gpu.rs:575-589 single query, :758-776 per query
 * of a batch): every row's score, in row order, pushed into one accumulator.
 * Where rows tie at the k-th score this keeps a different set from the
 * (score desc, row asc) order; the set is pinned by a closed-form restatement
 * of the heap's behaviour in tests/test_oracle.py. */
size_t orc_gpu_exact_search(const float *matrix, size_t rows, size_t dims, const float *query,
                            size_t k, uint64_t *idx, float *sim) {
    size_t capped = k < rows ? k : rows;
    orc_topk *acc = orc_topk_new(capped);
    for (size_t r = 0; r < rows; ++r)
        orc_topk_push(acc, r, orc_dot_product(matrix + r * dims, query, dims));
    return orc_topk_into_sorted(acc, idx, sim);
}

/* matrixmultiply sgemm restated: per (q, r) a sequential FMA chain over k
 * inside each KC=256 block; block partials are added into C in order. */
void orc_sgemm_scores(const float *queries, size_t query_rows, const float *matrix, size_t rows,
                      size_t dims, int threads, float *scores) {
    const size_t KC = 256;
    threads = resolve_threads(threads);
#pragma omp parallel for num_threads(threads) schedule(static) collapse(2)
    for (size_t q = 0; q < query_rows; ++q) {
        for (size_t r = 0; r < rows; ++r) {
            const float *a = queries + q * dims, *b = matrix + r * dims;
            float cval = 0.f;
            for (size_t k0 = 0; k0 < dims; k0 += KC) {
                size_t k1 = k0 + KC < dims ? k0 + KC : dims;
                float ab = 0.f;
                for (size_t kk = k0; kk < k1; ++kk) ab = fmaf(a[kk], b[kk], ab);
                cval = k0 == 0 ? ab : cval + ab;
            }
            scores[q * rows + r] = cval;
        }
    }
}

void orc_cpu_search_batch(const float *matrix, size_t rows, size_t dims, const float *queries,
                          size_t query_rows, size_t k, int threads, uint64_t *idx, float *sim,
                          size_t *out_lens, int *strategy_out) {
    cpu_ctx c = {matrix, rows, dims, resolve_threads(threads)};
    size_t capped = k < rows ? k : rows;
    int prefer_qp = query_rows >= (size_t)c.threads * 2;
    int s = orc_cpu_batch_strategy(rows, dims, query_rows, prefer_qp);
    if (strategy_out) *strategy_out = s;
    size_t batch_cutoff = orc_profile_cutoff(rows, dims, query_rows);

    if (s == ORC_B_GEMM) { /* cpu.rs:256-284 */
        float *scores = (float *)malloc(query_rows * rows * sizeof(float));
        orc_sgemm_scores(queries, query_rows, matrix, rows, dims, c.threads, scores);
#pragma omp parallel for num_threads(c.threads) schedule(dynamic, 1)
        for (size_t q = 0; q < query_rows; ++q)
            out_lens[q] =
                orc_topk_from_scores(scores + q * rows, rows, capped, 0, idx + q * capped,
                                     sim + q * capped);
        free(scores);
        return;
    }
    if (s == ORC_B_GEMM_CHUNKED || s == ORC_B_MATRIX_CHUNKED) {
        /* GemmChunked (cpu.rs:291-341): 64 MiB chunks, serial streaming merge.
         * MatrixChunked (cpu.rs:349-396): 256 KiB chunks, dot_product scores. */
        size_t chunk_rows = s == ORC_B_GEMM_CHUNKED ? (64u * 1024 * 1024) / (dims * 4)
                                                     : orc_batch_chunk_rows(dims);
        if (chunk_rows < 64) chunk_rows = 64;
        orc_topk **accs = (orc_topk **)malloc(query_rows * sizeof(orc_topk *));
        for (size_t q = 0; q < query_rows; ++q) accs[q] = orc_topk_new(capped);
        float *scores = (float *)malloc(query_rows * chunk_rows * sizeof(float));
        for (size_t r0 = 0; r0 < rows; r0 += chunk_rows) {
            size_t nr = r0 + chunk_rows < rows ? chunk_rows : rows - r0;
            if (s == ORC_B_GEMM_CHUNKED) {
                orc_sgemm_scores(queries, query_rows, matrix + r0 * dims, nr, dims, c.threads,
                                 scores);
            } else {
#pragma omp parallel for num_threads(c.threads) schedule(static) collapse(2)
                for (size_t q = 0; q < query_rows; ++q)
                    for (size_t r = 0; r < nr; ++r)
                        scores[q * nr + r] =
                            orc_dot_product(matrix + (r0 + r) * dims, queries + q * dims, dims);
            }
#pragma omp parallel for num_threads(c.threads) schedule(static)
            for (size_t q = 0; q < query_rows; ++q) {
                orc_topk *acc = accs[q];
                float threshold = acc->min_threshold;
                for (size_t r = 0; r < nr; ++r) {
                    float sc = scores[q * nr + r];
                    if (sc > threshold) {
                        orc_topk_push_slow(acc, r0 + r, sc);
                        threshold = acc->min_threshold;
                    }
                }
            }
        }
        free(scores);
        for (size_t q = 0; q < query_rows; ++q)
            out_lens[q] = orc_topk_into_sorted(accs[q], idx + q * capped, sim + q * capped);
        free(accs);
        return;
    }
    /* QueryParallel (cpu.rs:236-240) / ProfiledPerQuery (cpu.rs:241-245) */
    cpu_ctx inner = c;
    if (s == ORC_B_QUERY_PARALLEL) {
#pragma omp parallel for num_threads(c.threads) schedule(dynamic, 1)
        for (size_t q = 0; q < query_rows; ++q)
            out_lens[q] = orc_topk_into_sorted(search_serial(&inner, queries + q * dims, capped),
                                               idx + q * capped, sim + q * capped);
    } else {
        for (size_t q = 0; q < query_rows; ++q)
            out_lens[q] = orc_topk_into_sorted(
                search_with_profile(&inner, queries + q * dims, capped, batch_cutoff, NULL),
                idx + q * capped, sim + q * capped);
    }
}

/* ------------------------------------------------------------------------- */
/* FP64 brute-force oracle                                                   */
/* ------------------------------------------------------------------------- */

typedef struct {
    uint64_t index;
    double score;
} item64_t;

static int item64_cmp(const void *pa, const void *pb) {
    const item64_t *a = (const item64_t *)pa, *b = (const item64_t *)pb;
    if (a->score != b->score) return a->score > b->score ? -1 : 1;
    if (a->index != b->index) return a->index < b->index ? -1 : 1;
    return 0;
}

static int fp64_unit(const float *src, size_t dims, int normalized, double *dst) {
    double ss = 0.0;
    for (size_t i = 0; i < dims; ++i) {
        if (!isfinite(src[i])) return ORC_NON_FINITE;
        dst[i] = (double)src[i];
        ss += dst[i] * dst[i];
    }
    if (!normalized) {
        if (!(ss > 0.0)) return ORC_ZERO_NORM;
        double inv = 1.0 / sqrt(ss);
        for (size_t i = 0; i < dims; ++i) dst[i] *= inv;
    }
    return ORC_OK;
}

int orc_fp64_scores(const float *matrix, size_t rows, size_t dims, const float *query,
                    int normalized, int threads, double *scores) {
    threads = resolve_threads(threads);
    double *q = (double *)malloc(dims * sizeof(double));
    int st = fp64_unit(query, dims, normalized, q);
    if (st != ORC_OK) {
        free(q);
        return st;
    }
    int bad = ORC_OK;
#pragma omp parallel num_threads(threads)
    {
        double *row = (double *)malloc(dims * sizeof(double));
#pragma omp for schedule(static)
        for (size_t r = 0; r < rows; ++r) {
            int s = fp64_unit(matrix + r * dims, dims, normalized, row);
            if (s != ORC_OK) {
#pragma omp atomic write
                bad = s;
                scores[r] = NAN;
                continue;
            }
            double acc = 0.0;
            for (size_t i = 0; i < dims; ++i) acc += row[i] * q[i];
            scores[r] = acc;
        }
        free(row);
    }
    free(q);
    return bad;
}

int orc_fp64_topk(const float *matrix, size_t rows, size_t dims, const float *queries,
                  size_t query_rows, size_t k, int normalized, int threads, uint64_t *idx,
                  double *sim) {
    size_t capped = k < rows ? k : rows;
    double *scores = (double *)malloc(rows * sizeof(double));
    item64_t *items = (item64_t *)malloc(rows * sizeof(item64_t));
    for (size_t q = 0; q < query_rows; ++q) {
        int st = orc_fp64_scores(matrix, rows, dims, queries + q * dims, normalized, threads,
                                 scores);
        if (st != ORC_OK) {
            free(scores);
            free(items);
            return st;
        }
        for (size_t r = 0; r < rows; ++r) {
            items[r].index = r;
            items[r].score = scores[r];
        }
        qsort(items, rows, sizeof(item64_t), item64_cmp);
        for (size_t i = 0; i < capped; ++i) {
            idx[q * capped + i] = items[i].index;
            sim[q * capped + i] = items[i].score;
        }
    }
    free(scores);
    free(items);
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* IVF (approx.rs)                                                           */
/* ------------------------------------------------------------------------- */

#define MIN_ROWS_FOR_IVF 256

/* isqrt, approx.rs:339-350 */
size_t orc_isqrt(size_t n) {
    if (n == 0) return 0;
    size_t x = n, y = (x + 1) / 2;
    while (y < x) {
        x = y;
        y = (x + n / x) / 2;
    }
    return x;
}

/* choose_nlist, approx.rs:305-311 */
size_t orc_choose_nlist(size_t rows) {
    if (rows < MIN_ROWS_FOR_IVF) return 1;
    size_t s = orc_isqrt(rows);
    return s < 16 ? 16 : (s > 512 ? 512 : s);
}

/* choose_nprobe, approx.rs:318-336 */
size_t orc_choose_nprobe(size_t nlist, size_t k, size_t rows) {
    size_t pct = k >= 50 ? 8 : (k >= 20 ? 6 : 4);
    size_t nprobe = (nlist * pct + 99) / 100;
    if (nprobe < 1) nprobe = 1;
    size_t avg = rows / (nlist ? nlist : 1);
    if (avg < 1) avg = 1;
    size_t minc = k * 40;
    size_t min_nprobe = (minc + avg - 1) / avg;
    if (min_nprobe < 1) min_nprobe = 1;
    if (nprobe < min_nprobe) nprobe = min_nprobe;
    if (nprobe > nlist) nprobe = nlist;
    if (nprobe < 1) nprobe = 1;
    return nprobe;
}

struct orc_ivf {
    size_t rows, dims, nlist;
    float *centroids;
    float *reordered;
    uint64_t *original_indices;
    uint64_t *offsets;
    uint64_t *assignments;
};

/* pick_medoid_row, approx.rs:409-429 */
static size_t pick_medoid_row(const float *m, size_t rows, size_t dims) {
    double *mean = (double *)calloc(dims, sizeof(double));
    for (size_t r = 0; r < rows; ++r)
        for (size_t d = 0; d < dims; ++d) mean[d] += (double)m[r * dims + d];
    double inv = 1.0 / (double)rows;
    float *mf = (float *)malloc(dims * sizeof(float));
    for (size_t d = 0; d < dims; ++d) mf[d] = (float)(mean[d] * inv);
    size_t best = 0;
    float best_dot = -INFINITY;
    for (size_t r = 0; r < rows; ++r) {
        float d = dot_small(m + r * dims, mf, dims);
        if (d > best_dot) {
            best_dot = d;
            best = r;
        }
    }
    free(mean);
    free(mf);
    return best;
}

/* kmeans_plusplus_init, approx.rs:357-406 (distance pass parallel, sums serial).
 * chosen (may be NULL): the row each centroid copies (a duplicate repeats the
 * previous centroid's row). */
static float *kmeans_pp(const float *m, size_t rows, size_t dims, size_t nlist, int threads,
                        uint64_t *chosen_rows) {
    uint64_t state = (uint64_t)rows * 2654435761ull + (uint64_t)dims;
    float *cent = (float *)malloc(nlist * dims * sizeof(float));
    size_t first = pick_medoid_row(m, rows, dims);
    memcpy(cent, m + first * dims, dims * sizeof(float));
    if (chosen_rows) chosen_rows[0] = first;
    float *mind = (float *)malloc(rows * sizeof(float));
    for (size_t r = 0; r < rows; ++r) mind[r] = FLT_MAX;
    for (size_t c = 1; c < nlist; ++c) {
        const float *last = cent + (c - 1) * dims;
#pragma omp parallel for num_threads(threads) schedule(static)
        for (size_t r = 0; r < rows; ++r) {
            float d = 1.0f - dot_small(m + r * dims, last, dims);
            if (d < 0.0f) d = 0.0f;
            if (d < mind[r]) mind[r] = d;
        }
        double total = 0.0;
        for (size_t r = 0; r < rows; ++r) total += (double)mind[r];
        if (total <= 0.0) {
            memcpy(cent + c * dims, last, dims * sizeof(float));
            if (chosen_rows) chosen_rows[c] = chosen_rows[c - 1];
            continue;
        }
        state = state * 6364136223846793005ull + 1ull;
        double threshold = ((double)state / (double)UINT64_MAX) * total;
        double cum = 0.0;
        size_t chosen = rows - 1;
        for (size_t r = 0; r < rows; ++r) {
            cum += (double)mind[r];
            if (cum >= threshold) {
                chosen = r;
                break;
            }
        }
        memcpy(cent + c * dims, m + chosen * dims, dims * sizeof(float));
        if (chosen_rows) chosen_rows[c] = chosen;
    }
    free(mind);
    return cent;
}

size_t orc_pick_medoid_row(const float *matrix, size_t rows, size_t dims) {
    return rows ? pick_medoid_row(matrix, rows, dims) : 0;
}

void orc_kmeans_pp_rows(const float *matrix, size_t rows, size_t dims, size_t nlist, int threads,
                        uint64_t *chosen_rows) {
    if (!rows || !nlist) return;
    free(kmeans_pp(matrix, rows, dims, nlist, resolve_threads(threads), chosen_rows));
}

/* spherical_kmeans, approx.rs:434-535 */
static void spherical_kmeans(const float *m, size_t rows, size_t dims, size_t nlist, size_t iters,
                             float *cent, uint64_t *assign, int threads) {
    double *sums = (double *)malloc(nlist * dims * sizeof(double));
    uint64_t *sizes = (uint64_t *)malloc(nlist * sizeof(uint64_t));
    for (size_t it = 0; it < iters; ++it) {
#pragma omp parallel for num_threads(threads) schedule(static)
        for (size_t r = 0; r < rows; ++r) {
            size_t best = 0;
            float best_score = -INFINITY;
            for (size_t c = 0; c < nlist; ++c) {
                float s = dot_small(m + r * dims, cent + c * dims, dims);
                if (s > best_score) {
                    best_score = s;
                    best = c;
                }
            }
            assign[r] = best;
        }
        memset(sums, 0, nlist * dims * sizeof(double));
        memset(sizes, 0, nlist * sizeof(uint64_t));
        for (size_t r = 0; r < rows; ++r) {
            size_t c = assign[r];
            sizes[c] += 1;
            for (size_t d = 0; d < dims; ++d) sums[c * dims + d] += (double)m[r * dims + d];
        }
        for (size_t c = 0; c < nlist; ++c) {
            float *dst = cent + c * dims;
            if (sizes[c] == 0) {
                for (size_t d = 0; d < dims; ++d) dst[d] = 0.0f;
                continue;
            }
            double inv_count = 1.0 / (double)sizes[c];
            double nsq = 0.0;
            for (size_t d = 0; d < dims; ++d) {
                double v = sums[c * dims + d] * inv_count;
                nsq += v * v;
            }
            double inv_norm = nsq > 1e-30 ? 1.0 / sqrt(nsq) : 1.0;
            for (size_t d = 0; d < dims; ++d)
                dst[d] = (float)(sums[c * dims + d] * inv_count * inv_norm);
        }
    }
    free(sums);
    free(sizes);
}

/* ApproxIndex::build, approx.rs:46-100 */
orc_ivf *orc_ivf_build(const float *matrix, size_t rows, size_t dims, size_t nlist, size_t iters,
                       int threads) {
    threads = resolve_threads(threads);
    orc_ivf *ivf = (orc_ivf *)calloc(1, sizeof(orc_ivf));
    ivf->rows = rows;
    ivf->dims = dims;
    if (nlist == 0) nlist = orc_choose_nlist(rows);
    if (iters == 0) iters = 12;
    ivf->reordered = (float *)malloc(rows * dims * sizeof(float));
    ivf->original_indices = (uint64_t *)malloc(rows * sizeof(uint64_t));
    ivf->assignments = (uint64_t *)calloc(rows ? rows : 1, sizeof(uint64_t));
    if (rows < MIN_ROWS_FOR_IVF || nlist < 2) {
        ivf->nlist = 1;
        memcpy(ivf->reordered, matrix, rows * dims * sizeof(float));
        for (size_t r = 0; r < rows; ++r) ivf->original_indices[r] = r;
        ivf->offsets = (uint64_t *)malloc(2 * sizeof(uint64_t));
        ivf->offsets[0] = 0;
        ivf->offsets[1] = rows;
        return ivf;
    }
    ivf->nlist = nlist;
    ivf->centroids = kmeans_pp(matrix, rows, dims, nlist, threads, NULL);
    spherical_kmeans(matrix, rows, dims, nlist, iters, ivf->centroids, ivf->assignments, threads);
    ivf->offsets = (uint64_t *)calloc(nlist + 1, sizeof(uint64_t));
    for (size_t r = 0; r < rows; ++r) ivf->offsets[ivf->assignments[r] + 1] += 1;
    for (size_t c = 0; c < nlist; ++c) ivf->offsets[c + 1] += ivf->offsets[c];
    uint64_t *wp = (uint64_t *)malloc(nlist * sizeof(uint64_t));
    memcpy(wp, ivf->offsets, nlist * sizeof(uint64_t));
    for (size_t r = 0; r < rows; ++r) {
        uint64_t dst = wp[ivf->assignments[r]]++;
        ivf->original_indices[dst] = r;
        memcpy(ivf->reordered + dst * dims, matrix + r * dims, dims * sizeof(float));
    }
    free(wp);
    return ivf;
}

void orc_ivf_free(orc_ivf *ivf) {
    if (!ivf) return;
    free(ivf->centroids);
    free(ivf->reordered);
    free(ivf->original_indices);
    free(ivf->offsets);
    free(ivf->assignments);
    free(ivf);
}

size_t orc_ivf_nlist(const orc_ivf *ivf) { return ivf->nlist; }
int orc_ivf_provides_speedup(const orc_ivf *ivf) {
    return ivf->rows >= MIN_ROWS_FOR_IVF && ivf->nlist >= 4;
}
const float *orc_ivf_centroids(const orc_ivf *ivf) { return ivf->centroids; }
const uint64_t *orc_ivf_offsets(const orc_ivf *ivf) { return ivf->offsets; }
const uint64_t *orc_ivf_original_indices(const orc_ivf *ivf) { return ivf->original_indices; }
const uint64_t *orc_ivf_assignments(const orc_ivf *ivf) { return ivf->assignments; }

/* candidate_clusters, approx.rs:159-215: centroid scores (ndarray GEMV when
 * nlist>=16 && dims>=16, else dot_small), top nprobe by score desc.
 * select_nth_unstable_by has no index tie-break; ties resolve to lower id here. */
/* Top-nprobe clusters by centroid score (score desc, id asc).  gemv != 0
 * scores with ndarray's dot (candidate_clusters, approx.rs:158-196, used by
 * the CPU backend); gemv == 0 with dot_small (candidate_indices,
 * approx.rs:110-138, used by the GPU backend). */
static size_t ivf_probe_impl(const orc_ivf *ivf, const float *query, size_t nprobe,
                             uint64_t *clusters, int allow_gemv) {
    size_t nl = ivf->nlist;
    item_t *sc = (item_t *)malloc(nl * sizeof(item_t));
    int gemv = allow_gemv && nl >= 16 && ivf->dims >= 16;
    for (size_t c = 0; c < nl; ++c) {
        const float *cen = ivf->centroids + c * ivf->dims;
        sc[c].index = c;
        sc[c].similarity =
            gemv ? orc_dot_ndarray(cen, query, ivf->dims) : dot_small(cen, query, ivf->dims);
    }
    qsort(sc, nl, sizeof(item_t), out_cmp);
    if (nprobe > nl) nprobe = nl;
    for (size_t i = 0; i < nprobe; ++i) clusters[i] = sc[i].index;
    free(sc);
    return nprobe;
}

size_t orc_ivf_probe(const orc_ivf *ivf, const float *query, size_t nprobe, uint64_t *clusters) {
    return ivf_probe_impl(ivf, query, nprobe, clusters, 1);
}

/* search_approx_cpu of the GPU backend, src/gpu.rs:448-462: candidate_indices
 * (approx.rs:110-138: dot_small centroid scores, top nprobe), then every
 * candidate row of the original matrix rescored with cpu_dot
 * (src/gpu.rs:887-898, the simsimd dot for len >= 8) into a TopKAccumulator. */
size_t orc_ivf_search_gpu(const orc_ivf *ivf, const float *matrix, const float *query, size_t k,
                          size_t nprobe, uint64_t *idx, float *sim) {
    size_t dims = ivf->dims;
    size_t capped = k < ivf->rows ? k : ivf->rows;
    orc_topk *acc = orc_topk_new(capped);
    if (nprobe == 0) nprobe = orc_choose_nprobe(ivf->nlist, capped, ivf->rows);
    uint64_t *cl = (uint64_t *)malloc((ivf->nlist ? ivf->nlist : 1) * sizeof(uint64_t));
    size_t np = ivf->nlist <= 1 ? 1 : ivf_probe_impl(ivf, query, nprobe, cl, 0);
    if (ivf->nlist <= 1) cl[0] = 0;
    for (size_t p = 0; p < np; ++p) {
        size_t s = ivf->offsets[cl[p]], e = ivf->offsets[cl[p] + 1];
        for (size_t r = s; r < e; ++r) {
            uint64_t row = ivf->original_indices[r];
            orc_topk_push(acc, row, orc_dot_product(matrix + row * dims, query, dims));
        }
    }
    free(cl);
    return orc_topk_into_sorted(acc, idx, sim);
}

/* search_approx, cpu.rs:425-453 */
size_t orc_ivf_search(const orc_ivf *ivf, const float *query, size_t k, size_t nprobe,
                      uint64_t *idx, float *sim) {
    size_t dims = ivf->dims;
    size_t capped = k < ivf->rows ? k : ivf->rows;
    orc_topk *acc = orc_topk_new(capped);
    if (ivf->nlist <= 1) {
        int use_gemv = ivf->rows >= 4 && dims >= 8;
        for (size_t r = 0; r < ivf->rows; ++r) {
            const float *row = ivf->reordered + r * dims;
            orc_topk_push(acc, ivf->original_indices[r],
                          use_gemv ? orc_dot_ndarray(row, query, dims)
                                   : orc_dot_product(row, query, dims));
        }
        return orc_topk_into_sorted(acc, idx, sim);
    }
    if (nprobe == 0) nprobe = orc_choose_nprobe(ivf->nlist, capped, ivf->rows);
    uint64_t *cl = (uint64_t *)malloc(ivf->nlist * sizeof(uint64_t));
    size_t np = orc_ivf_probe(ivf, query, nprobe, cl);
    for (size_t p = 0; p < np; ++p) {
        size_t s = ivf->offsets[cl[p]], e = ivf->offsets[cl[p] + 1];
        size_t n = e - s;
        int use_gemv = n >= 4 && dims >= 8;
        for (size_t r = s; r < e; ++r) {
            const float *row = ivf->reordered + r * dims;
            float score = use_gemv ? orc_dot_ndarray(row, query, dims)
                                   : orc_dot_product(row, query, dims);
            orc_topk_push(acc, ivf->original_indices[r], score);
        }
    }
    free(cl);
    return orc_topk_into_sorted(acc, idx, sim);
}
