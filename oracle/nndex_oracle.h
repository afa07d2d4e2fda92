/*
 * nndex_oracle.h -- CPU restatement of nndex's exact search path, used ONLY as
 * test infrastructure: the parity checker for the HIP backend and the timed CPU
 * baseline beside it. Nothing in the product (nndex_amd/) links or calls this.
 *
 * Every function names the reference file:line it restates (paths relative to
 * the fblissjr/nndex v0.2.1 checkout).  Where the reference delegates arithmetic
 * to a third-party crate that is not vendored (simsimd ^6.5.12, ndarray ^0.16.1
 * via matrixmultiply), the published algorithm of that crate is restated and the
 * header of nndex_oracle.c says which results are pinned bit-for-bit and which
 * only to a tolerance.
 */
#ifndef NNDEX_ORACLE_H
#define NNDEX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes, the subset of NNdexError (src/lib.rs:107-145) the CPU path raises. */
enum {
    ORC_OK = 0,
    ORC_NON_FINITE = 1, /* NNdexError::NonFiniteInput */
    ORC_ZERO_NORM = 2,  /* NNdexError::ZeroNorm */
};

/* Strategy ids, src/profiles.rs:30-63. */
enum {
    ORC_S_SERIAL = 0,
    ORC_S_GEMV = 1,
    ORC_S_PARALLEL_SCORES = 2,
    ORC_S_PARALLEL_CHUNKED = 3,
    ORC_S_PARALLEL_FOLD = 4,
};
enum {
    ORC_B_GEMM = 0,
    ORC_B_GEMM_CHUNKED = 1,
    ORC_B_MATRIX_CHUNKED = 2,
    ORC_B_QUERY_PARALLEL = 3,
    ORC_B_PROFILED_PER_QUERY = 4,
};

/* ---- normalisation and padding (src/lib.rs:433-599) ---- */
int orc_ensure_finite(const float *values, size_t n);
void orc_pad_rows(const float *values, size_t rows, size_t dims, size_t padded_dims, float *out);
int orc_normalize_rows(const float *values, size_t rows, size_t dims, size_t padded_dims,
                       float *out);
int orc_normalize_vector(const float *values, size_t dims, size_t padded_dims, float *out);

/* ---- dot products (src/cpu.rs:651-695, simsimd dot_f32) ---- */
float orc_dot_product(const float *lhs, const float *rhs, size_t n);
float orc_dot_product_simd(const float *lhs, const float *rhs, size_t n);
float orc_dot_unrolled_8(const float *lhs, const float *rhs, size_t n);
float orc_dot_ndarray(const float *lhs, const float *rhs, size_t n);

/* ---- top-k (src/topk.rs:44-210) ---- */
typedef struct orc_topk orc_topk;
orc_topk *orc_topk_new(size_t k);
void orc_topk_free(orc_topk *acc);
void orc_topk_push(orc_topk *acc, uint64_t index, float similarity);
void orc_topk_push_slow(orc_topk *acc, uint64_t index, float similarity);
void orc_topk_merge(orc_topk *acc, orc_topk *other); /* consumes `other` */
float orc_topk_min_threshold(const orc_topk *acc);
size_t orc_topk_len(const orc_topk *acc);
/* consumes `acc`; writes at most k entries, returns the count */
size_t orc_topk_into_sorted(orc_topk *acc, uint64_t *idx, float *sim);
size_t orc_topk_from_scores(const float *scores, size_t n, size_t k, size_t row_offset,
                            uint64_t *idx, float *sim);

/* ---- strategy selection (src/profiles.rs:72-185, src/cpu.rs:697-721) ---- */
size_t orc_profile_cutoff(size_t rows, size_t dims, size_t query_rows);
int orc_cpu_search_strategy(size_t profile_cutoff, size_t rows, size_t dims,
                            size_t parallel_cutoff_rows, size_t chunk_rows,
                            size_t scores_vec_max_rows, int has_blas);
int orc_cpu_batch_strategy(size_t rows, size_t dims, size_t query_rows, int prefer_query_parallel);
size_t orc_parallel_cutoff(size_t dims, int threads);
size_t orc_chunk_rows(size_t rows, size_t dims);
size_t orc_batch_chunk_rows(size_t dims);

/* ---- exact CPU index (src/cpu.rs:93-639) over normalised, padded storage ----
 * threads <= 0 means "all cores" (rayon's default pool size). */
size_t orc_cpu_search(const float *matrix, size_t rows, size_t dims, const float *query,
                      size_t k, int threads, uint64_t *idx, float *sim, int *strategy_out);
void orc_cpu_search_batch(const float *matrix, size_t rows, size_t dims, const float *queries,
                          size_t query_rows, size_t k, int threads, uint64_t *idx, float *sim,
                          size_t *out_lens, int *strategy_out);
/* GPU backend's exact search (src/gpu.rs:575-589): rows pushed in order into
 * one accumulator. */
size_t orc_gpu_exact_search(const float *matrix, size_t rows, size_t dims, const float *query,
                            size_t k, uint64_t *idx, float *sim);
/* sgemm restatement: scores[q*rows + r] = Q[q,:] . X[r,:] */
void orc_sgemm_scores(const float *queries, size_t query_rows, const float *matrix, size_t rows,
                      size_t dims, int threads, float *scores);

/* ---- FP64 brute-force oracle ----
 * Inputs are the RAW caller values.  normalized=0: rows and queries are L2
 * normalised in f64; normalized=1: used as given.  Output is sorted by
 * (score desc, index asc) in f64. */
int orc_fp64_topk(const float *matrix, size_t rows, size_t dims, const float *queries,
                  size_t query_rows, size_t k, int normalized, int threads, uint64_t *idx,
                  double *sim);
int orc_fp64_scores(const float *matrix, size_t rows, size_t dims, const float *query,
                    int normalized, int threads, double *scores);

/* ---- IVF (src/approx.rs:1-552) ---- */
size_t orc_isqrt(size_t n);
size_t orc_choose_nlist(size_t rows);
size_t orc_choose_nprobe(size_t nlist, size_t k, size_t rows);

/* pick_medoid_row (approx.rs:409-429): the row with the largest dot with the
 * f64-summed column mean (rounded to f32); the lowest row on ties. */
size_t orc_pick_medoid_row(const float *matrix, size_t rows, size_t dims);
/* kmeans_plusplus_init (approx.rs:357-406): the row each of the nlist seeded
 * centroids copies (a duplicate, when every distance is zero, repeats the
 * previous centroid's row). */
void orc_kmeans_pp_rows(const float *matrix, size_t rows, size_t dims, size_t nlist, int threads,
                        uint64_t *chosen_rows);

typedef struct orc_ivf orc_ivf;
/* nlist == 0 -> choose_nlist(rows); iters == 0 -> 12 (KMEANS_MAX_ITERS) */
orc_ivf *orc_ivf_build(const float *matrix, size_t rows, size_t dims, size_t nlist, size_t iters,
                       int threads);
void orc_ivf_free(orc_ivf *ivf);
size_t orc_ivf_nlist(const orc_ivf *ivf);
int orc_ivf_provides_speedup(const orc_ivf *ivf);
const float *orc_ivf_centroids(const orc_ivf *ivf);
const uint64_t *orc_ivf_offsets(const orc_ivf *ivf);
const uint64_t *orc_ivf_original_indices(const orc_ivf *ivf);
const uint64_t *orc_ivf_assignments(const orc_ivf *ivf);
/* nprobe == 0 -> choose_nprobe(nlist, k, rows) */
size_t orc_ivf_search(const orc_ivf *ivf, const float *query, size_t k, size_t nprobe,
                      uint64_t *idx, float *sim);
/* probed cluster ids of one query, sorted by (score desc, id asc); returns count */
size_t orc_ivf_probe(const orc_ivf *ivf, const float *query, size_t nprobe, uint64_t *clusters);
/* GPU-backend approximate search (src/gpu.rs:448-462) over `matrix`, the rows
 * the index was built from; nprobe == 0 -> choose_nprobe(nlist, k, rows) */
size_t orc_ivf_search_gpu(const orc_ivf *ivf, const float *matrix, const float *query, size_t k,
                          size_t nprobe, uint64_t *idx, float *sim);

#ifdef __cplusplus
}
#endif

#endif
