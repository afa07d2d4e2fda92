// Kernel launch interfaces shared by the backend translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nndex {

// Launch-shape flags of the fused exact scan.
constexpr uint32_t kScanNonTemporal = 1;  // streaming (non-temporal) loads
constexpr uint32_t kScanContiguous = 2;   // contiguous tile runs per wave
// bits 2-3: load group of the single-query scan, 0 = 8, 1 = 12, 2 = 16 float4
constexpr uint32_t kScanGroupShift = 2;

// Arguments of the fused exact scan (exact.hip).
struct ScanArgs {
    const float4 *x;              // tiled matrix, rows padded to 64
    uint32_t rows;                // logical rows
    uint32_t d4;                  // padded_dims / 4
    const float4 *q;              // QB queries, row-major, stride q_stride4 float4
    uint32_t q_stride4;
    uint32_t nvalid;              // queries actually present (<= QB); later ones repeat the last
    uint32_t k;                   // keys selected per query (1 <= k <= rows)
    uint32_t k_out;               // neighbours written per query (<= k): when k >= 2 k_out - 1
                                  // the reference heap's choice among ties (heap_tie_shift)
    uint64_t *part;               // [QB][grid][k] each block's published list (key 0: empty)
    uint32_t *part_count;         // [QB] not used by the scan (kept 0)
    unsigned long long *theta;    // [QB] not used by the scan (kept 0)
    uint32_t *ticket;             // block-completion counter (0 between launches)
    uint64_t *out_idx;            // [QB][out_stride]
    float *out_sim;               // [QB][out_stride]
    uint32_t out_stride;
    uint32_t flags;               // kScan* launch-shape flags
};

void launch_to_tiled(const float *src, uint32_t rows, uint32_t dims, uint32_t d4, float4 *dst,
                     uint32_t *flags, hipStream_t s);
void launch_normalize_tiled(float4 *x, uint32_t rows, uint32_t dims, uint32_t d4,
                            uint32_t *flags, hipStream_t s);
// Query preparation modes of launch_normalize_rowmajor.
constexpr int kQueryPad = 0;     // ensure_finite + pad (normalized = true)
constexpr int kQueryBatch = 1;   // normalize_rows (batch queries)
constexpr int kQuerySingle = 2;  // normalize_vector (a single query)
// Rows failing the checks are zeroed; flag bits (1 non-finite, 2 zero norm)
// are ORed into *flags and, if row_flags is given, stored per row.
void launch_normalize_rowmajor(const float *src, uint32_t rows, uint32_t dims, uint32_t pdims,
                               int mode, float *dst, uint32_t *flags, uint8_t *row_flags,
                               hipStream_t s);

// Slots per wave for a given k (power of two >= k + 64); 0 QB means k is too
// large for the fused scan.
uint32_t scan_slots_for_k(uint32_t k);
uint32_t scan_max_qb(uint32_t k);
void launch_exact_scan(const ScanArgs &a, uint32_t qb, uint32_t grid, hipStream_t s);
// Plain 16-byte streaming copy (bandwidth calibration).
void launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s);

}  // namespace nndex
