// Launch interfaces of the batched MFMA search (batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nndex {

// Element type of the MFMA scan (the candidate-scoring precision).
enum ElemKind : int { kElemF32 = 0, kElemBF16 = 1, kElemFP8 = 2 };

struct MfmaScanArgs {
    const uint4 *x;            // tiled matrix in 16-byte units (f32: 4 values, bf16: 8, fp8: 16)
    uint32_t rows;             // logical rows
    uint32_t rows_padded;      // rows rounded up to 64 (allocated tiles)
    uint32_t dch;              // 16-byte units per row
    const uint4 *q;            // queries, row-major, q_stride units per query
    uint32_t q_stride;
    const float *row_scale;    // fp8: per-row scale [rows_padded]
    const float *q_scale;      // fp8: per-query scale [nq]
    uint32_t nq;
    uint32_t mode;             // 0: append candidates, 1: write scores, 2: row argmax,
                               // 3: write each group's largest score (out[nq][ldo])
    const unsigned long long *thresh;  // [nq] key thresholds (register staging)
    const float *thresh_f;     // [nq] their float floors (mode 0, padded: see qimg)
    uint64_t *cand;            // [nq][capq] candidate keys (mode 0)
    uint32_t *cand_count;      // [nq] (mode 0; may exceed capq on overflow)
    uint32_t capq;
    float *out;                // [nq][ldo] scores (mode 1)
    uint32_t ldo;
    // mode 2: per row, the best query (ties -> lowest id); a row whose best
    // and second-best approximate scores are within *tol is appended to
    // `recheck` for an exact re-evaluation
    uint32_t *assign;          // [rows]
    const float *tol;          // device scalar
    uint32_t *recheck;         // [rows]
    uint32_t *recheck_count;
    // query splits (grid y): each block row scans its share of the query
    // chunks (0 or 1: one share)
    uint32_t qsplit;
    // modes 0 and 1 without qsplit: the row tiles of a last, partial round
    // of the grid are split into this many pieces of the query chunks
    // (0 or 1: not split)
    uint32_t tail_split;
    // the batch's chunk image (launch_chunk_image for this element type,
    // width and variant, or the argmax shape), which the scan copies into LDS;
    // `thresh_f` (mode 0) and `q_scale` (FP8) are then read a whole chunk at
    // a time, so both hold nq rounded up to the chunk's query count (the
    // image's padding entries set by launch_chunk_image)
    const uint4 *qimg;
};

struct RerankArgs {
    const float4 *x;           // row-major FP32 copy of the master rows (d4 float4 per row)
    uint32_t d4;
    const float4 *q;           // FP32 queries, row-major, d4 float4 each
    const uint64_t *cand;
    const uint32_t *cand_count;
    uint32_t capq;
    const unsigned long long *thresh;
    const float *delta;
    uint32_t k;                // results per query
    uint32_t kp;               // rescored candidates per query (k' >= k)
    uint64_t *out_idx;         // [nq][k]
    float *out_sim;            // [nq][k]
    uint8_t *flags;            // [nq] 1 = not certified
};

void launch_to_bf16(const float4 *x, size_t rows_padded, uint32_t d8, uint4 *y, hipStream_t s);
// FP8-E4M3 (OCP) copy with a per-row scale s_r = max|x_r| / 448 (at least
// 2^-126; its exponent is the MFMA's E8M0 block scale): tiled units of
// 16 codes (d16 = ceil(d4 / 4) units per row), scales [rows_padded], and the
// largest per-row norm of |s_r x~_r - x_r| (f64 bits, atomic max into *err_sq).
void launch_to_fp8(const float4 *x, size_t rows_padded, uint32_t d4, uint4 *y, float *scale,
                   unsigned long long *err_sq, hipStream_t s);
void launch_max_norm(const float4 *x, uint32_t rows, uint32_t d4, unsigned long long *out,
                     hipStream_t s);
void launch_prep_queries(const float *q, uint32_t nq, uint32_t pdims, int elem, uint4 *qlow,
                         float *q_scale, const unsigned long long *xmax_sq,
                         const unsigned long long *xerr_sq, float *delta, hipStream_t s);
void launch_bf16_error(const float4 *x, uint32_t rows, uint32_t d4, unsigned long long *out,
                       hipStream_t s);
// Tile-shape variants of the candidate / score scan for a row width (1 or 2;
// the autotuner times them), rows per block of a variant (the argmax scan
// has its own tiles), and the launch.
uint32_t mfma_variants(int elem, uint32_t dch);
uint32_t mfma_block_rows(int elem, uint32_t dch, bool argmax = false, int variant = 0);
void launch_mfma_scan(const MfmaScanArgs &a, int elem, uint32_t grid, hipStream_t s,
                      int variant = 0);
// Chunk shape of the candidate / score scan (or, with argmax, of the row
// argmax) for (element type, units per row, variant): queries per chunk and
// unit rows per query (4 * steps >= dch).
void mfma_chunk_shape(int elem, uint32_t dch, int variant, uint32_t &qc, uint32_t &ru,
                      bool argmax = false);
// Largest chunk image of any shape for nq queries, in 16-byte units.
inline size_t mfma_image_units(size_t nq) { return (nq + 64) * 128; }
// Writes the chunk image of nq row-major queries (q_stride units apart, dch
// valid units each) for the scan shape (qc, ru), and sets the padding
// entries of thresh_pad (+inf) and scale_pad (0) up to the last chunk;
// either may be null.
void launch_chunk_image(const uint4 *q, uint32_t q_stride, uint32_t nq, uint32_t dch, uint32_t qc,
                        uint32_t ru, uint4 *img, float *thresh_pad, float *scale_pad,
                        hipStream_t s);
// Per query the key threshold (the kp-th best sample score, row part zero)
// and its float floor (key_floor_sim), which the candidate scan compares.
void launch_sample_threshold(const float *scores, uint32_t nq, uint32_t n, uint32_t ld,
                             uint32_t kp, const float *delta, unsigned long long *thresh,
                             float *thresh_f, hipStream_t s);
void launch_rerank(const RerankArgs &a, uint32_t nq, hipStream_t s);
void launch_rerank_full(const RerankArgs &a, uint32_t nq, hipStream_t s);
// Tiled FP32 rows -> row-major (d4 float4 per row): the reranks read whole
// candidate rows, which the tiled layout spreads over d4 separate 1 KiB lines.
void launch_tiled_to_rows(const float4 *x, uint32_t rows, uint32_t d4, float4 *y, hipStream_t s);
void launch_gather_rows(const float4 *x, uint32_t d4, uint32_t stride, uint32_t rows_s,
                        float4 *dst, hipStream_t s);
void launch_zero_u32(uint32_t *p, size_t n, hipStream_t s);
// One wavefront's chain of `calls` MFMAs of element form `elem` over device
// operands (per call and lane: 4 bytes FP32, 16 bytes BF16, 32 bytes FP8),
// the 16 x 16 FP32 result row-major into out (device).
void launch_mfma_tile(int elem, const void *a, const void *b, uint32_t calls, float *out,
                      hipStream_t s);
// Row-major FP32 rows -> row-major BF16 (16-byte units of 8), and the largest
// per-row norm of the rounding error (f64 bits, atomic max into *err_sq).
void launch_rows_to_bf16(const float *rows, uint32_t n, uint32_t pdims, uint4 *out,
                         unsigned long long *err_sq, hipStream_t s);

}  // namespace nndex
