// IVF (inverted file) prefilter on the GPU: spherical k-means build, cluster-
// contiguous reorder, probe and exact scan of the probed lists.
//
// Reference: src/approx.rs (ApproxIndex::build :46-100, kmeans_plusplus_init
// :357-406, pick_medoid_row :409-429, spherical_kmeans :434-535,
// candidate_indices :110-138) as the reference GPU backend uses it
// (src/gpu.rs:435-462: built lazily from the padded rows, probed with the
// simsimd dot, candidates rescored with the same dot into a top-k).
//
// Build, on the device:
//   * medoid: f64 column means, then the row with the largest dot with the
//     mean (strict >, the lowest row wins ties);
//   * k-means++ seeding with the reference's rule and generator: distance
//     max(0, 1 - dot) to the newest centroid, running minimum per row, f64
//     totals, LCG state = state * 6364136223846793005 + 1 seeded with
//     rows * 2654435761 + dims, first row whose cumulative distance reaches
//     state / u64::MAX * total, duplicate of the last centroid when total <= 0;
//   * Lloyd iterations: assignment = MFMA GEMM of rows x centroids (FP32 up
//     to 512 padded dims, BF16 copies above) with
//     each row's two best centroids; a row whose two best approximate scores
//     lie within the rounding bound is re-evaluated in the reference dot
//     order, so the argmax equals the reference's (lowest id on ties);
//     update = stable radix sort of (cluster, row), then one block per cluster
//     sums its rows in row order in f64, takes the mean and normalises
//     (empty cluster -> zero centroid);
//   * reorder: rows in cluster order (row order inside a cluster) into a
//     second tiled FP32 copy plus the reordered -> original map.
// Search: centroid scores and the top nprobe clusters through the exact path
// of a centroid index, then each probed list scanned in the reference dot
// order into a per-list top-k, merged per query.  Every score is the exact
// FP32 value the reference computes.
#include <cstring>  // rocPRIM's headers use memset

#include <algorithm>
#include <chrono>
#include <mutex>
#include <rocprim/device/device_radix_sort.hpp>
#include <vector>

#include "batch.hpp"
#include "common.hpp"
#include "gpu_index.hpp"
#include "ivf.hpp"
#include "select.hpp"
#include "wave_topk.hpp"

namespace nndex {

// Reference dot of a tiled row with a row-major vector, in the simsimd
// four-lane order (src/approx.rs:538-551 dot_small for len >= 8).
__device__ inline float tiled_dot(const float4 *__restrict__ x, uint32_t row, uint32_t d4,
                                  const float4 *__restrict__ v) {
    const float4 *b = x + (size_t)(row / 64) * d4 * 64 + row % 64;
    float l0 = 0.f, l1 = 0.f, l2 = 0.f, l3 = 0.f;
    // unrolled: the row's loads are requested together, not one per step
#pragma unroll 8
    for (uint32_t c = 0; c < d4; ++c) {
        const float4 a = b[(size_t)c * 64], q = v[c];
        l0 = __builtin_fmaf(a.x, q.x, l0);
        l1 = __builtin_fmaf(a.y, q.y, l1);
        l2 = __builtin_fmaf(a.z, q.z, l2);
        l3 = __builtin_fmaf(a.w, q.w, l3);
    }
    return (l0 + l1) + (l2 + l3);
}

// Key for an argmax under IEEE `>` with the first (lowest) id kept on ties:
// -0.0 and +0.0 compare equal there, so both map to +0.0 (x + 0.0f).
__device__ inline uint64_t argmax_key(float s, uint32_t id) { return make_key(s + 0.0f, id); }

// Maximum over the block (every thread receives it); red has a slot per wave.
__device__ inline uint64_t block_max_u64(uint64_t v, uint64_t *red) {
    for (int o = 32; o > 0; o >>= 1) {
        const uint64_t w = __shfl_xor(v, o);
        v = w > v ? w : v;
    }
    const uint32_t wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (__lane_id() == 0) red[wid] = v;
    __syncthreads();
    uint64_t m = 0;
    for (uint32_t i = 0; i < nw; ++i) m = red[i] > m ? red[i] : m;
    return m;
}

// ---- medoid (pick_medoid_row, approx.rs:409-429) ---------------------------

// Per-block f64 column sums over a contiguous row range: partial[b][dim].
__global__ __launch_bounds__(256) void colsum_kernel(const float4 *__restrict__ x, uint32_t rows,
                                                     uint32_t d4, uint32_t rows_per_block,
                                                     double *__restrict__ partial) {
    const uint32_t r0 = blockIdx.x * rows_per_block;
    const uint32_t r1 = r0 + rows_per_block < rows ? r0 + rows_per_block : rows;
    for (uint32_t c = threadIdx.x; c < d4; c += blockDim.x) {  // float4 chunk
        double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll 4
        for (uint32_t r = r0; r < r1; ++r) {
            const float4 v = x[((size_t)(r / 64) * d4 + c) * 64 + r % 64];
            s0 += v.x;
            s1 += v.y;
            s2 += v.z;
            s3 += v.w;
        }
        double *p = partial + (size_t)blockIdx.x * d4 * 4 + (size_t)c * 4;
        p[0] = s0;
        p[1] = s1;
        p[2] = s2;
        p[3] = s3;
    }
}

__global__ __launch_bounds__(256) void colmean_kernel(const double *__restrict__ partial,
                                                      uint32_t nblocks, uint32_t pdims,
                                                      uint32_t rows, float *__restrict__ mean) {
    for (uint32_t d = threadIdx.x; d < pdims; d += 256) {
        double s = 0;
        for (uint32_t b = 0; b < nblocks; ++b) s += partial[(size_t)b * pdims + d];
        mean[d] = (float)(s * (1.0 / (double)rows));
    }
}

// Row argmax of the reference dot with v (lowest row on ties).
__global__ __launch_bounds__(256) void dot_argmax_kernel(const float4 *__restrict__ x,
                                                         uint32_t rows, uint32_t d4,
                                                         const float4 *__restrict__ v,
                                                         unsigned long long *__restrict__ best) {
    const uint32_t row = blockIdx.x * 256 + threadIdx.x;
    unsigned long long key = 0;
    if (row < rows) key = argmax_key(tiled_dot(x, row, d4, v), row);
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long other = __shfl_xor(key, o);
        key = other > key ? other : key;
    }
    if (__lane_id() == 0) atomicMax(best, key);
}

__device__ inline void copy_row(const float4 *__restrict__ x, uint32_t row, uint32_t d4,
                                float4 *__restrict__ dst) {
    for (uint32_t c = threadIdx.x; c < d4; c += blockDim.x)
        dst[c] = x[((size_t)(row / 64) * d4 + c) * 64 + row % 64];
}

__global__ void copy_key_row_kernel(const float4 *__restrict__ x, uint32_t d4,
                                    const unsigned long long *__restrict__ key,
                                    float4 *__restrict__ dst) {
    copy_row(x, key_row(*key), d4, dst);
}

// ---- k-means++ seeding (kmeans_plusplus_init, approx.rs:357-406) ------------

// The new centroid's values rounded to BF16 (as the BF16 row copy rounds).
__global__ void bf16_round_row_kernel(const float4 *__restrict__ v, uint32_t d4,
                                      float4 *__restrict__ out) {
    for (uint32_t c = threadIdx.x; c < d4; c += blockDim.x) {
        const float4 a = v[c];
        auto r = [](float f) { return (float)(__bf16)f; };
        out[c] = make_float4(r(a.x), r(a.y), r(a.z), r(a.w));
    }
}

// min_dist[r] = min(min_dist[r], max(0, 1 - dot(row, last))); f64 block sums.
//
// With the BF16 row copy (xb, d8 units of 8 per row; lastb the centroid
// rounded to BF16) a row is first scored approximately: a = x~.c~ in f32 is
// within delta of x.c, and the reference's f32 value 1 - x.c within eps of
// exact, so (1 - a) - (delta + eps) > m proves `d < m` false and the FP32 row
// is not read (margin = delta + eps, from the largest row norm and BF16
// rounding error).  Other rows take the reference's computation.
constexpr uint32_t kKppRows = 256;    // rows per k-means++ update block
constexpr uint32_t kKppBlock = 1024;  // threads of the choice kernel

// Sum over a block of NT threads (one value per thread); thread 0 gets the
// total.
template <uint32_t NT>
__device__ inline double kpp_block_sum(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (uint32_t i = 0; i < NT / 64; ++i) t += red[i];
    return t;
}

// Inclusive prefix sum over the block of kKppBlock threads (wave scans, then
// the wave totals in order).  `red` holds kKppBlock / 64 doubles; the call
// ends with a barrier, so `red` may be reused at once.
__device__ inline double kpp_block_scan(double v, double *red) {
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const double u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    if (lane == 63) red[w] = v;
    __syncthreads();
    double off = 0.0;
    for (uint32_t i = 0; i < w; ++i) off += red[i];
    __syncthreads();
    return off + v;
}

__global__ __launch_bounds__(kKppRows) void kpp_update_kernel(const float4 *__restrict__ x,
                                                         const uint4 *__restrict__ xb,
                                                         uint32_t rows, uint32_t d4,
                                                         const float4 *__restrict__ last,
                                                         const float4 *__restrict__ lastb,
                                                         double margin,
                                                         float *__restrict__ min_dist,
                                                         double *__restrict__ partial) {
    __shared__ double red[kKppRows / 64];
    const uint32_t row = blockIdx.x * kKppRows + threadIdx.x;
    double md = 0.0;
    if (row < rows) {
        float m = min_dist[row];
        bool skip = false;
        if (xb && m < 3.0e38f) {
            const uint32_t d8 = d4 / 2;
            const uint4 *b = xb + (size_t)(row / 64) * d8 * 64 + row % 64;
            const_as<f32x4> *cb = as_constant(reinterpret_cast<const f32x4 *>(lastb));
            float l[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll 8
            for (uint32_t c = 0; c < d8; ++c) {
                const uint4 w = b[(size_t)c * 64];
                const f32x4 q0 = cb[2 * c], q1 = cb[2 * c + 1];
                const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
                const float qs[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float lo = __builtin_bit_cast(float, ws[j] << 16);
                    const float hi = __builtin_bit_cast(float, ws[j] & 0xFFFF0000u);
                    l[2 * j] = __builtin_fmaf(lo, qs[2 * j], l[2 * j]);
                    l[2 * j + 1] = __builtin_fmaf(hi, qs[2 * j + 1], l[2 * j + 1]);
                }
            }
            const float a = ((l[0] + l[1]) + (l[2] + l[3])) + ((l[4] + l[5]) + (l[6] + l[7]));
            skip = (1.0 - (double)a) - margin > (double)m * (1.0 + 1e-12) + 1e-30;
        }
        if (!skip) {
            float d = 1.0f - tiled_dot(x, row, d4, last);
            if (d < 0.0f) d = 0.0f;
            if (d < m) {
                m = d;
                min_dist[row] = m;
            }
        }
        md = (double)m;
    }
    const double t = kpp_block_sum<kKppRows>(md, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// The same pass with the pre-test on the FP8 copy (half the bytes of the
// BF16 one): a = s_r * sum(code_i * c_i) over eight FMA chains (codes decoded
// by v_cvt_pk_f32_fp8, OCP E4M3), against the FP32 centroid; |a - x.c| <=
// gamma |x~||c| + |x~ - x||c|, folded with the reference value's own error
// into `margin` by the caller.
__global__ __launch_bounds__(kKppRows) void kpp_update_fp8_kernel(
    const float4 *__restrict__ x, const uint4 *__restrict__ x8, const float *__restrict__ xs,
    uint32_t rows, uint32_t d4, const float4 *__restrict__ last, double margin,
    float *__restrict__ min_dist, double *__restrict__ partial) {
    typedef float v2f __attribute__((ext_vector_type(2)));
    __shared__ double red[kKppRows / 64];
    const uint32_t row = blockIdx.x * kKppRows + threadIdx.x;
    double md = 0.0;
    if (row < rows) {
        float m = min_dist[row];
        bool skip = false;
        if (m < 3.0e38f) {
            const uint32_t d16 = (d4 + 3) / 4;
            const uint4 *b = x8 + (size_t)(row / 64) * d16 * 64 + row % 64;
            const_as<f32x4> *cb = as_constant(reinterpret_cast<const f32x4 *>(last));
            float l[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll 8
            for (uint32_t u = 0; u < d16; ++u) {
                const uint4 w = b[(size_t)u * 64];
                const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t c4 = 4 * u + j;  // the float4 these four codes cover
                    if (c4 >= d4) break;            // (wave-uniform) padding codes
                    const f32x4 q = cb[c4];
                    const v2f lo = __builtin_amdgcn_cvt_pk_f32_fp8((int)ws[j], false);
                    const v2f hi = __builtin_amdgcn_cvt_pk_f32_fp8((int)ws[j], true);
                    l[2 * j] = __builtin_fmaf(lo.x, q.x, l[2 * j]);
                    l[2 * j] = __builtin_fmaf(lo.y, q.y, l[2 * j]);
                    l[2 * j + 1] = __builtin_fmaf(hi.x, q.z, l[2 * j + 1]);
                    l[2 * j + 1] = __builtin_fmaf(hi.y, q.w, l[2 * j + 1]);
                }
            }
            const float a =
                xs[row] * (((l[0] + l[1]) + (l[2] + l[3])) + ((l[4] + l[5]) + (l[6] + l[7])));
            skip = (1.0 - (double)a) - margin > (double)m * (1.0 + 1e-12) + 1e-30;
        }
        if (!skip) {
            float d = 1.0f - tiled_dot(x, row, d4, last);
            if (d < 0.0f) d = 0.0f;
            if (d < m) {
                m = d;
                min_dist[row] = m;
            }
        }
        md = (double)m;
    }
    const double t = kpp_block_sum<kKppRows>(md, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// One block of kKppBlock threads: total, generator step, then the first row
// whose cumulative distance reaches the target, found top-down (ranges of
// update blocks, update blocks, rows) with block-wide prefix sums.  Writes
// centroid c and, when `lastb` is set, its BF16 rounding for the next update
// pass.
__global__ __launch_bounds__(kKppBlock) void kpp_choose_kernel(
    const float4 *__restrict__ x, uint32_t rows, uint32_t d4, const float *__restrict__ min_dist,
    const double *__restrict__ partial, uint32_t nblocks, unsigned long long *__restrict__ rng,
    float4 *__restrict__ centroids, uint32_t c, float4 *__restrict__ lastb) {
    __shared__ double red[kKppBlock / 64];
    __shared__ double base, target;
    __shared__ int dup;
    __shared__ uint32_t first, chosen;
    const uint32_t t = threadIdx.x;
    // this thread's range of update blocks, summed in order (loads batched)
    const uint32_t per = (nblocks + kKppBlock - 1) / kKppBlock;
    const uint32_t b0 = t * per < nblocks ? t * per : nblocks;
    const uint32_t b1 = b0 + per < nblocks ? b0 + per : nblocks;
    double own = 0.0;
    {
        uint32_t b = b0;
        for (; b + 8 <= b1; b += 8) {
            double v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = partial[b + j];
#pragma unroll
            for (int j = 0; j < 8; ++j) own += v[j];
        }
        for (; b < b1; ++b) own += partial[b];
    }
    if (t == 0) first = 0xFFFFFFFFu;
    double incl = kpp_block_scan(own, red);
    if (t == kKppBlock - 1) {
        const double total = incl;
        dup = !(total > 0.0);
        if (!dup) {
            const unsigned long long st = *rng * 6364136223846793005ull + 1ull;
            *rng = st;
            target = ((double)st / (double)0xFFFFFFFFFFFFFFFFull) * total;
        }
        chosen = rows - 1;  // rounding fallback
    }
    __syncthreads();
    float4 *dst = centroids + (size_t)c * d4;
    if (dup) {
        for (uint32_t k = t; k < d4; k += blockDim.x) dst[k] = centroids[(size_t)(c - 1) * d4 + k];
        return;  // lastb already holds that centroid's rounding
    }
    if (incl >= target) atomicMin(&first, t);
    __syncthreads();
    const uint32_t sel = first;
    // the last thread's scan value is the total, so some thread is found
    const uint32_t r0b = sel < kKppBlock ? sel * per : nblocks;
    if (r0b < nblocks) {
        // the update block holding the crossing, kKppBlock block sums at a time
        if (t == sel) base = incl - own;
        const uint32_t r1b = r0b + per < nblocks ? r0b + per : nblocks;
        uint32_t blk = r1b - 1;  // rounding fallback: the range's last block
        for (uint32_t j0 = r0b; j0 < r1b; j0 += kKppBlock) {
            __syncthreads();
            const double start = base;
            if (t == 0) first = 0xFFFFFFFFu;
            const double v = j0 + t < r1b ? partial[j0 + t] : 0.0;
            const double cum = start + kpp_block_scan(v, red);
            if (j0 + t < r1b && cum >= target) atomicMin(&first, t);
            __syncthreads();
            const uint32_t f = first;
            if (f != 0xFFFFFFFFu) {
                if (t == f) base = cum - v;
                blk = j0 + f;
                break;
            }
            if (t == kKppBlock - 1) base = cum;
        }
        __syncthreads();  // base written, every read of `first` done
        if (t == 0) first = 0xFFFFFFFFu;
        // the row inside that block
        const uint32_t r0 = blk * kKppRows, n = rows - r0 < kKppRows ? rows - r0 : kKppRows;
        const double start = base;
        const double v = t < n ? (double)min_dist[r0 + t] : 0.0;
        const double cum = start + kpp_block_scan(v, red);
        if (t < n && cum >= target) atomicMin(&first, t);
        __syncthreads();
        if (t == 0) chosen = first != 0xFFFFFFFFu ? r0 + first : r0 + n - 1;
    }
    __syncthreads();
    copy_row(x, chosen, d4, dst);
    if (lastb) {
        const uint32_t row = chosen;
        for (uint32_t k = t; k < d4; k += blockDim.x) {
            const float4 a = x[((size_t)(row / 64) * d4 + k) * 64 + row % 64];
            auto r = [](float f) { return (float)(__bf16)f; };
            lastb[k] = make_float4(r(a.x), r(a.y), r(a.z), r(a.w));
        }
    }
}

__global__ void fill_f32_kernel(float *p, size_t n, float v) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        p[i] = v;
}

__global__ void iota_kernel(uint32_t *p, uint32_t n) {
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) p[i] = i;
}

// Cluster sizes: per-block LDS histogram, then one global add per cluster.
__global__ __launch_bounds__(256) void histogram_kernel(const uint32_t *__restrict__ assign,
                                                        uint32_t rows, uint32_t nlist,
                                                        uint32_t *__restrict__ counts) {
    extern __shared__ uint32_t hist[];
    for (uint32_t i = threadIdx.x; i < nlist; i += 256) hist[i] = 0;
    __syncthreads();
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < rows; i += gridDim.x * 256)
        atomicAdd(hist + assign[i], 1u);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nlist; i += 256)
        if (hist[i]) atomicAdd(counts + i, hist[i]);
}

// Exclusive scan of nlist counts into offsets[nlist + 1] (one block).
__global__ __launch_bounds__(256) void offsets_kernel(const uint32_t *__restrict__ counts,
                                                      uint32_t nlist,
                                                      uint32_t *__restrict__ offsets) {
    __shared__ uint32_t part[256];
    const uint32_t per = (nlist + 255) / 256, t = threadIdx.x;
    const uint32_t i0 = t * per, i1 = (t + 1) * per < nlist ? (t + 1) * per : nlist;
    uint32_t s = 0;
    for (uint32_t i = i0; i < i1; ++i) s += counts[i];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        uint32_t acc = 0;
        for (int i = 0; i < 256; ++i) {
            const uint32_t v = part[i];
            part[i] = acc;
            acc += v;
        }
    }
    __syncthreads();
    uint32_t acc = part[t];
    for (uint32_t i = i0; i < i1; ++i) {
        offsets[i] = acc;
        acc += counts[i];
    }
    if (t == 255) offsets[nlist] = acc;
}

// Centroid update (spherical_kmeans, approx.rs:490-531): one block per
// cluster; each thread owns float4 chunks t, t + blockDim, ... and sums the
// cluster's rows in row order in f64; the squared norm of the mean is summed
// in dimension order.  Dynamic LDS: 4 d4 doubles.
__global__ __launch_bounds__(1024) void centroid_update_kernel(const float4 *__restrict__ x,
                                                               uint32_t d4,
                                                               const uint32_t *__restrict__ sorted_rows,
                                                               const uint32_t *__restrict__ offsets,
                                                               float4 *__restrict__ centroids) {
    extern __shared__ double scaled[];
    __shared__ double nsq;
    const uint32_t cl = blockIdx.x;
    const uint32_t r0 = offsets[cl], r1 = offsets[cl + 1];
    const uint32_t count = r1 - r0;
    const double inv_count = count ? 1.0 / (double)count : 0.0;
    for (uint32_t t = threadIdx.x; t < d4; t += blockDim.x) {
        double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll 4
        for (uint32_t i = r0; i < r1; ++i) {
            const uint32_t r = sorted_rows[i];
            const float4 v = x[((size_t)(r / 64) * d4 + t) * 64 + r % 64];
            s0 += v.x;
            s1 += v.y;
            s2 += v.z;
            s3 += v.w;
        }
        scaled[4 * t + 0] = s0 * inv_count;
        scaled[4 * t + 1] = s1 * inv_count;
        scaled[4 * t + 2] = s2 * inv_count;
        scaled[4 * t + 3] = s3 * inv_count;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double acc = 0;
        for (uint32_t i = 0; i < 4 * d4; ++i) acc += scaled[i] * scaled[i];
        nsq = acc;
    }
    __syncthreads();
    const double inv_norm = nsq > 1e-30 ? 1.0 / sqrt(nsq) : 1.0;
    for (uint32_t t = threadIdx.x; t < d4; t += blockDim.x) {
        float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
        if (count) {
            // (sum / count) / norm: the sums scaled once, as the reference
            out = make_float4((float)(scaled[4 * t + 0] * inv_norm),
                              (float)(scaled[4 * t + 1] * inv_norm),
                              (float)(scaled[4 * t + 2] * inv_norm),
                              (float)(scaled[4 * t + 3] * inv_norm));
        }
        centroids[(size_t)cl * d4 + t] = out;
    }
}

// Exact argmax for the rows the MFMA pass could not separate: the reference
// dot (four-lane order) of the row with every centroid, lowest id on ties.
__global__ __launch_bounds__(256) void recheck_kernel(const float4 *__restrict__ x, uint32_t d4,
                                                      const float4 *__restrict__ centroids,
                                                      uint32_t nlist,
                                                      const uint32_t *__restrict__ list,
                                                      const uint32_t *__restrict__ count,
                                                      uint32_t *__restrict__ assign) {
    __shared__ float4 srow[256];
    __shared__ uint64_t red[4];
    const uint32_t n = *count;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const uint32_t row = list[i];
        __syncthreads();
        for (uint32_t c = threadIdx.x; c < d4; c += 256)
            srow[c] = x[((size_t)(row / 64) * d4 + c) * 64 + row % 64];
        __syncthreads();
        uint64_t best = 0;
        for (uint32_t cl = threadIdx.x; cl < nlist; cl += 256) {
            const float4 *cv = centroids + (size_t)cl * d4;
            float l0 = 0.f, l1 = 0.f, l2 = 0.f, l3 = 0.f;
#pragma unroll 8
            for (uint32_t c = 0; c < d4; ++c) {
                const float4 a = srow[c], b = cv[c];
                l0 = __builtin_fmaf(a.x, b.x, l0);
                l1 = __builtin_fmaf(a.y, b.y, l1);
                l2 = __builtin_fmaf(a.z, b.z, l2);
                l3 = __builtin_fmaf(a.w, b.w, l3);
            }
            const uint64_t key = argmax_key((l0 + l1) + (l2 + l3), cl);
            best = key > best ? key : best;
        }
        best = block_max_u64(best, red);
        if (threadIdx.x == 0) assign[row] = key_row(best);
    }
}

// Exact row argmax for rows too wide for the MFMA assignment (padded dims
// above 1024): a tile of 64 rows against every centroid, 64 centroids at a
// time, chunks of 8 float4 of both staged in LDS.  Thread (ty, tx) = (t / 16,
// t % 16) owns rows ty + 16 i and centroids tx + 16 j (i, j < 4) and keeps,
// per pair, the reference dot's four FMA chains in registers across chunks
// (lane j accumulates elements 4c + j in chunk order, then (l0 + l1) +
// (l2 + l3)), so every score is the reference's f32 value and the argmax
// needs no re-check: the highest score, the lowest centroid on ties (strict
// >, approx.rs:447-469).
constexpr uint32_t kWideChunk = 8;
__global__ __launch_bounds__(256) void assign_exact_kernel(const float4 *__restrict__ x,
                                                           uint32_t rows, uint32_t d4,
                                                           const float4 *__restrict__ centroids,
                                                           uint32_t nlist,
                                                           uint32_t *__restrict__ assign) {
    __shared__ float4 xs[kWideChunk][64];
    __shared__ float4 cs[kWideChunk][64];
    const uint32_t t = threadIdx.x, ty = t / 16, tx = t % 16;
    const uint32_t tile = blockIdx.x;
    uint64_t best[4] = {0, 0, 0, 0};
    for (uint32_t c0 = 0; c0 < nlist; c0 += 64) {
        float acc[4][4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j][0] = acc[i][j][1] = acc[i][j][2] = acc[i][j][3] = 0.f;
        for (uint32_t k0 = 0; k0 < d4; k0 += kWideChunk) {
            const uint32_t kn = d4 - k0 < kWideChunk ? d4 - k0 : kWideChunk;
            __syncthreads();
            for (uint32_t u = t; u < kWideChunk * 64; u += 256) {
                const uint32_t kc = u / 64, r = u % 64;
                float4 xv = make_float4(0.f, 0.f, 0.f, 0.f), cv = xv;
                if (kc < kn) {
                    // tiled rows: chunk k0 + kc of the tile's 64 rows (padding
                    // rows of the last tile are zero)
                    xv = x[((size_t)tile * d4 + k0 + kc) * 64 + r];
                    if (c0 + r < nlist) cv = centroids[(size_t)(c0 + r) * d4 + k0 + kc];
                }
                xs[kc][r] = xv;
                cs[kc][r] = cv;
            }
            __syncthreads();
            for (uint32_t kc = 0; kc < kn; ++kc) {
                float4 xa[4], cb[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) xa[i] = xs[kc][ty + 16 * i];
#pragma unroll
                for (int j = 0; j < 4; ++j) cb[j] = cs[kc][tx + 16 * j];
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        acc[i][j][0] = __builtin_fmaf(xa[i].x, cb[j].x, acc[i][j][0]);
                        acc[i][j][1] = __builtin_fmaf(xa[i].y, cb[j].y, acc[i][j][1]);
                        acc[i][j][2] = __builtin_fmaf(xa[i].z, cb[j].z, acc[i][j][2]);
                        acc[i][j][3] = __builtin_fmaf(xa[i].w, cb[j].w, acc[i][j][3]);
                    }
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t cid = c0 + tx + 16 * j;
                if (cid < nlist) {
                    const float sc = (acc[i][j][0] + acc[i][j][1]) + (acc[i][j][2] + acc[i][j][3]);
                    const uint64_t key = argmax_key(sc, cid);
                    best[i] = key > best[i] ? key : best[i];
                }
            }
    }
    // the 16 threads of a row group (consecutive lanes) hold its candidates
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        uint64_t b = best[i];
        for (int off = 1; off < 16; off <<= 1) {
            const uint64_t o = __shfl_xor(b, off);
            b = o > b ? o : b;
        }
        const uint32_t row = tile * 64 + ty + 16 * i;
        if (tx == 0 && row < rows) assign[row] = key_row(b);
    }
}

// Re-check margin of the assignment.  The approximate (MFMA) score and the
// reference dot lie within da and de of the exact dot:
//   de = gamma_n |x| |c|                                  (four-lane FMA chain)
//   da = gamma_n |x~| |c~| + |x~ - x| |c~| + |x| |c~ - c|  (x~, c~ = BF16 copies;
//        for FP32 x~ = x, c~ = c and only the first term remains)
// so a top-two gap above 2 (da + de) fixes the reference's argmax; the margin
// used is twice that.  |x| <= x_max, |c| <= c_max, |x~ - x| <= x_err, and
// |c~ - c| is the measured largest centroid rounding error.
__global__ void assign_tol_kernel(const unsigned long long *__restrict__ cerr_sq, double x_err,
                                  double x_max, double c_max, uint32_t pdims, int bf16,
                                  float *__restrict__ tol) {
    const double u = 0x1p-24, n = pdims;
    const double gamma = n * u / (1.0 - n * u);
    const double de = gamma * x_max * c_max;
    double da = de;
    if (bf16) {
        const double ue = 0x1p-8;
        const double xl = x_max * (1.0 + ue), cl = c_max * (1.0 + ue);
        const double ec = sqrt(__builtin_bit_cast(double, *cerr_sq));
        da = gamma * xl * cl + x_err * cl + x_max * ec;
    }
    *tol = (float)(4.0 * (da + de) * (1.0 + 1e-6));
}

// Reorder: reordered row j = row sorted[j] (both tiled, padding rows zero).
__global__ __launch_bounds__(256) void gather_sorted_kernel(const float4 *__restrict__ x,
                                                            uint32_t rows, uint32_t d4,
                                                            const uint32_t *__restrict__ sorted,
                                                            float4 *__restrict__ y) {
    const size_t total = (size_t)((rows + 63) / 64) * 64 * d4;
    for (size_t u = (size_t)blockIdx.x * 256 + threadIdx.x; u < total;
         u += (size_t)gridDim.x * 256) {
        const size_t tile = u / ((size_t)d4 * 64), rem = u % ((size_t)d4 * 64);
        const uint32_t c = rem / 64, r = rem % 64;
        const size_t j = tile * 64 + r;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (j < rows) {
            const uint32_t src = sorted[j];
            v = x[((size_t)(src / 64) * d4 + c) * 64 + src % 64];
        }
        y[u] = v;
    }
}

// ---- search -----------------------------------------------------------------

// Pair i = q * nprobe + p -> (list, pair) for the sort by list.
__global__ void pair_prep_kernel(const uint64_t *__restrict__ probe, uint32_t pairs,
                                 uint32_t *__restrict__ key, uint32_t *__restrict__ val) {
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < pairs; i += gridDim.x * 256) {
        key[i] = (uint32_t)probe[i];
        val[i] = i;
    }
}

// Arrival base of each (query, probe): the rows of the query's lists probed
// before it (the reference pushes rows list by list in probe order).
__global__ void pair_base_kernel(const uint64_t *__restrict__ probe, uint32_t nq, uint32_t nprobe,
                                 const uint32_t *__restrict__ offsets,
                                 uint32_t *__restrict__ base, uint32_t *__restrict__ totals) {
    for (uint32_t q = blockIdx.x * 256 + threadIdx.x; q < nq; q += gridDim.x * 256) {
        uint32_t acc = 0;
        for (uint32_t p = 0; p < nprobe; ++p) {
            base[(size_t)q * nprobe + p] = acc;
            const uint32_t c = (uint32_t)probe[(size_t)q * nprobe + p];
            acc += offsets[c + 1] - offsets[c];
        }
        if (totals) totals[q] = acc;
    }
}

// Wavefronts per list: ceil(pairs of the list / qb).
__global__ void group_count_kernel(const uint32_t *__restrict__ pair_cnt, uint32_t nlist,
                                   uint32_t qb, uint32_t *__restrict__ grp_cnt) {
    for (uint32_t c = blockIdx.x * 256 + threadIdx.x; c < nlist; c += gridDim.x * 256)
        grp_cnt[c] = (pair_cnt[c] + qb - 1) / qb;
}

// Wavefront = (pair group of one list, row chunk of that list).  Lane r of a
// 64-row tile scores row r against the group's QB queries in the reference
// dot order (four-lane FMA chains, (l0 + l1) + (l2 + l3)); the queries come
// through the scalar cache; each query keeps its own wave top-k in LDS and
// writes its best kk keys for this chunk.  A key's row part is the row's
// arrival index in the query's probe sequence (list by list in probe order,
// reordered position within a list), the order the reference pushes rows in.
// KEYS (the large-k form): no top-k; every row's key goes to
// keys_out[query][arrival].
template <int QB, int CW, bool KEYS = false>
__global__ __launch_bounds__(256) void ivf_list_scan_kernel(IvfScanArgs a) {
    __shared__ uint64_t slots[4][QB][KEYS ? 1 : CW];
    // wave-uniform by construction; readfirstlane lets the compiler keep the
    // work item, the list and the query pointers in scalar registers
    const uint32_t lane = __lane_id();
    const uint32_t wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t item = blockIdx.x * 4 + wid;
    const uint32_t g = item / a.nchunk, chunk = item % a.nchunk;
    if (g >= a.grp_off[a.nlist]) return;  // wave-uniform; no block barrier below
    // the list: grp_off[c] <= g < grp_off[c + 1]
    uint32_t lo = 0, hi = a.nlist;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (a.grp_off[mid] <= g)
            lo = mid;
        else
            hi = mid;
    }
    const uint32_t c = lo;
    const uint32_t p0 = a.pair_off[c] + (g - a.grp_off[c]) * QB;
    const uint32_t pn = a.pair_off[c + 1] - p0 < (uint32_t)QB ? a.pair_off[c + 1] - p0 : QB;
    uint32_t pid[QB];
    const_as<f32x4> *Q[QB];
#pragma unroll
    for (int j = 0; j < QB; ++j) {
        pid[j] = a.sorted_pair[p0 + ((uint32_t)j < pn ? j : pn - 1)];
        Q[j] = as_constant(
            reinterpret_cast<const f32x4 *>(a.q + (size_t)(pid[j] / a.nprobe) * a.d4 * 4));
    }
    const uint32_t l0 = a.offsets[c], len = a.offsets[c + 1] - l0;
    const uint32_t per = (len + a.nchunk - 1) / a.nchunk;
    const uint32_t r0 = l0 + (chunk * per < len ? chunk * per : len);
    const uint32_t r1 = l0 + ((chunk + 1) * per < len ? (chunk + 1) * per : len);
    WaveTopK tk[QB];
    if constexpr (!KEYS) {
#pragma unroll
        for (int j = 0; j < QB; ++j) tk[j].init(slots[wid][j], CW, a.kk, nullptr, a.k);
    }
    uint32_t abase[QB];
#pragma unroll
    for (int j = 0; j < QB; ++j) abase[j] = a.pair_base[pid[j]] - l0;
    for (uint32_t t = r0 / 64; t * 64 < r1; ++t) {
        const uint32_t row = t * 64 + lane;
        const bool valid = row >= r0 && row < r1;
        const float4 *b = a.x + (size_t)t * a.d4 * 64 + lane;
        float acc[QB][4];
#pragma unroll
        for (int j = 0; j < QB; ++j) acc[j][0] = acc[j][1] = acc[j][2] = acc[j][3] = 0.f;
#pragma unroll 4
        for (uint32_t c4 = 0; c4 < a.d4; ++c4) {
            const float4 v = b[(size_t)c4 * 64];
#pragma unroll
            for (int j = 0; j < QB; ++j) {
                const f32x4 qv = Q[j][c4];
                acc[j][0] = __builtin_fmaf(v.x, qv.x, acc[j][0]);
                acc[j][1] = __builtin_fmaf(v.y, qv.y, acc[j][1]);
                acc[j][2] = __builtin_fmaf(v.z, qv.z, acc[j][2]);
                acc[j][3] = __builtin_fmaf(v.w, qv.w, acc[j][3]);
            }
        }
#pragma unroll
        for (int j = 0; j < QB; ++j) {
            if ((uint32_t)j >= pn) break;
            const float sc = (acc[j][0] + acc[j][1]) + (acc[j][2] + acc[j][3]);
            if constexpr (KEYS) {
                // the arrival index in 32 bits: abase wraps below zero when
                // the list starts past the query's base, and only the 32-bit
                // sum abase + row wraps back to the index
                const uint32_t arr = abase[j] + row;
                if (valid)
                    a.keys_out[(size_t)(pid[j] / a.nprobe) * a.keys_ld + arr] = make_key(sc, arr);
            } else {
                tk[j].offer(valid ? make_key(sc, abase[j] + row) : 0, valid);
            }
        }
    }
    if constexpr (KEYS) return;
#pragma unroll
    for (int j = 0; j < QB; ++j) {
        if ((uint32_t)j >= pn) break;
        tk[j].compact();
        uint64_t *dst = a.part + ((size_t)pid[j] * a.nchunk + chunk) * a.kk;
        for (uint32_t i = lane; i < a.kk; i += 64)
            dst[i] = i < tk[j].count ? slots[wid][j][i] : 0;
    }
}

// One block per query: its nprobe x nchunk partial lists merged into the
// final top-k and the number found.  The merged list is in (score desc,
// arrival asc) order; the reference heap's choice among rows tied at the k-th
// score follows from it (heap_tie_shift over arrival indices, then the heap
// evicts the lowest original rows first), and the results are written in
// (score desc, original row asc) order as the reference sorts them.
template <int CW>
__global__ __launch_bounds__(256) void ivf_merge_kernel(IvfScanArgs a) {
    __shared__ uint64_t slots[4][CW];
    __shared__ uint32_t cnt[4];
    const uint32_t q = blockIdx.x, lane = __lane_id(), wid = threadIdx.x >> 6;
    WaveTopK tk;
    tk.init(slots[wid], CW, a.kk, nullptr, a.k);
    const uint32_t n = a.nprobe * a.nchunk * a.kk;
    const uint64_t *src = a.part + (size_t)q * n;
    for (uint32_t i0 = wid * 64; i0 < n; i0 += 256) {
        const uint32_t i = i0 + lane;
        const uint64_t key = i < n ? src[i] : 0;
        tk.offer(key, key != 0);
    }
    tk.compact();
    if (lane == 0) cnt[wid] = tk.count;
    __syncthreads();
    if (wid != 0) return;
    for (uint32_t w = 1; w < 4; ++w)
        for (uint32_t i0 = 0; i0 < cnt[w]; i0 += 64) {
            const uint32_t i = i0 + lane;
            tk.offer(i < cnt[w] ? slots[w][i] : 0, i < cnt[w]);
        }
    tk.compact();
    const uint64_t *L = slots[0];
    const uint32_t m = tk.count < a.k ? tk.count : a.k;  // results
    uint32_t g = 0, e = 0;
    heap_tie_shift(L, tk.count, a.k, a.kk + 1 >= 2 * a.k, g, e);
    // window L[0, m + e): rows above the k-th score, then the tied rows the
    // heap saw before its k-th row >= that score; e of those are evicted
    const uint32_t w = m + e;
    uint32_t *ids = reinterpret_cast<uint32_t *>(slots[1]);  // waves 1-3 are done
    const uint32_t *pb = a.pair_base + (size_t)q * a.nprobe;
    for (uint32_t i = lane; i < w; i += 64) {
        const uint32_t arr = key_row(L[i]);
        uint32_t lo = 0, hi = a.nprobe;  // last probe whose base <= arr
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (pb[mid] <= arr)
                lo = mid;
            else
                hi = mid;
        }
        const uint32_t c = (uint32_t)a.probe[(size_t)q * a.nprobe + lo];
        ids[i] = a.orig[a.offsets[c] + (arr - pb[lo])];
    }
    wave_lds_sync();
    uint64_t *res = slots[2];
    for (uint32_t i = lane; i < w; i += 64) {
        bool keep = true;
        if (i >= g && e) {
            uint32_t rank = 0;  // among the tied window rows, by original row
            for (uint32_t j = g; j < w; ++j) rank += ids[j] < ids[i];
            keep = rank >= e;
        }
        res[i] = keep ? ((L[i] & 0xFFFFFFFF00000000ull) | (0xFFFFFFFFu - ids[i])) : 0;
    }
    wave_lds_sync();
    WaveTopK out;
    out.init(res, CW, w, nullptr);
    out.count = w;
    out.sort_desc();
    for (uint32_t i = lane; i < a.k; i += 64) {
        const uint64_t key = i < m ? res[i] : 0;
        a.out_idx[(size_t)q * a.k + i] = key ? key_row(key) : 0;
        a.out_sim[(size_t)q * a.k + i] = key ? key_sim(key) : 0.f;
    }
    if (lane == 0) a.out_len[q] = m;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

size_t ivf_isqrt(size_t n) {
    if (n == 0) return 0;
    size_t x = n, y = (x + 1) / 2;
    while (y < x) {
        x = y;
        y = (x + n / x) / 2;
    }
    return x;
}

size_t ivf_choose_nlist(size_t rows) {
    if (rows < kMinRowsForIvf) return 1;
    const size_t s = ivf_isqrt(rows);
    return s < 16 ? 16 : (s > 512 ? 512 : s);
}

size_t ivf_choose_nprobe(size_t nlist, size_t k, size_t rows) {
    const size_t pct = k >= 50 ? 8 : (k >= 20 ? 6 : 4);
    size_t nprobe = (nlist * pct + 99) / 100;
    if (nprobe < 1) nprobe = 1;
    size_t avg = rows / (nlist ? nlist : 1);
    if (avg < 1) avg = 1;
    size_t min_np = (k * 40 + avg - 1) / avg;
    if (min_np < 1) min_np = 1;
    if (nprobe < min_np) nprobe = min_np;
    if (nprobe > nlist) nprobe = nlist;
    return nprobe < 1 ? 1 : nprobe;
}

IvfIndex::~IvfIndex() {
    for (void *p : {(void *)centroids, (void *)offsets, (void *)orig, (void *)assign, (void *)xr})
        if (p) (void)hipFree(p);
}

static uint32_t ivf_cw(uint32_t k) {
    uint32_t cw = 128;
    while (cw < k + 64) cw <<= 1;
    return cw;
}

void ivf_build(IvfIndex &ivf, const float4 *x, const uint4 *xbf16, uint32_t rows,
               uint32_t pdims, uint32_t nlist, uint32_t iters, double x_max, double x_err,
               const Fp8Rows &fp8, uint32_t ncu, hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t d4 = pdims / 4, d8 = pdims / 8;
    // assignment: FP32 MFMA up to 512 padded dims, BF16 MFMA up to 1024 (both
    // with the exact re-check), the exact tiled kernel above
    const bool wide = pdims > 1024;
    const bool bf16 = pdims > 512 && !wide;
    if (pdims > 16384)
        throw Error(NNDEX_ERR_UNSUPPORTED, "the GPU IVF build supports padded dims <= 16384");
    if (bf16 && !xbf16)
        throw Error(NNDEX_ERR_INVALID_ARGUMENT, "the BF16 copy is required above 512 dims");
    if (nlist < 2 || nlist > rows || nlist > 16384)
        throw Error(NNDEX_ERR_INVALID_ARGUMENT, "nlist must be in [2, min(rows, 16384)]");
    ivf.rows = rows;
    ivf.pdims = pdims;
    ivf.nlist = nlist;
    ivf.iters = iters;
    const size_t prow = padded_rows(rows);
    NNDEX_HIP_CHECK(hipMalloc(&ivf.centroids, (size_t)nlist * pdims * sizeof(float)));
    NNDEX_HIP_CHECK(hipMalloc(&ivf.offsets, (nlist + 1) * sizeof(uint32_t)));
    NNDEX_HIP_CHECK(hipMalloc(&ivf.orig, (size_t)rows * sizeof(uint32_t)));
    NNDEX_HIP_CHECK(hipMalloc(&ivf.assign, (size_t)rows * sizeof(uint32_t)));
    NNDEX_HIP_CHECK(hipMalloc(&ivf.xr, prow * d4 * sizeof(float4)));

    const uint32_t nb = (rows + 255) / 256;
    const uint32_t cs_rows = 4096, cs_blocks = (rows + cs_rows - 1) / cs_rows;
    double *partial = nullptr, *colpart = nullptr;
    float *min_dist = nullptr, *mean = nullptr;
    float4 *lastb = nullptr;
    unsigned long long *rng = nullptr, *medoid = nullptr;
    uint32_t *sorted_keys = nullptr, *rows_iota = nullptr, *counts = nullptr, *recheck = nullptr,
             *recheck_count = nullptr;
    uint4 *cbf16 = nullptr, *cimg = nullptr;
    unsigned long long *cerr = nullptr;
    float *tol = nullptr;
    void *sort_tmp = nullptr;
    size_t sort_bytes = 0;
    auto free_all = [&] {
        for (void *p : {(void *)partial, (void *)colpart, (void *)min_dist, (void *)mean,
                        (void *)rng, (void *)medoid, (void *)sorted_keys, (void *)rows_iota,
                        (void *)counts, (void *)recheck, (void *)recheck_count, (void *)cbf16,
                        (void *)cimg,
                        (void *)cerr, (void *)tol, (void *)lastb, sort_tmp})
            if (p) (void)hipFree(p);
    };
    try {
        NNDEX_HIP_CHECK(hipMalloc(&partial, nb * sizeof(double)));
        NNDEX_HIP_CHECK(hipMalloc(&colpart, (size_t)cs_blocks * pdims * sizeof(double)));
        NNDEX_HIP_CHECK(hipMalloc(&min_dist, (size_t)rows * sizeof(float)));
        NNDEX_HIP_CHECK(hipMalloc(&mean, pdims * sizeof(float)));
        NNDEX_HIP_CHECK(hipMalloc(&rng, sizeof(unsigned long long)));
        NNDEX_HIP_CHECK(hipMalloc(&medoid, sizeof(unsigned long long)));
        NNDEX_HIP_CHECK(hipMalloc(&sorted_keys, (size_t)rows * sizeof(uint32_t)));
        NNDEX_HIP_CHECK(hipMalloc(&rows_iota, (size_t)rows * sizeof(uint32_t)));
        NNDEX_HIP_CHECK(hipMalloc(&counts, nlist * sizeof(uint32_t)));
        NNDEX_HIP_CHECK(hipMalloc(&recheck, (size_t)rows * sizeof(uint32_t)));
        NNDEX_HIP_CHECK(hipMalloc(&recheck_count, sizeof(uint32_t)));
        NNDEX_HIP_CHECK(hipMalloc(&cerr, 2 * sizeof(unsigned long long)));
        NNDEX_HIP_CHECK(hipMalloc(&tol, sizeof(float)));
        NNDEX_HIP_CHECK(hipMalloc(&lastb, d4 * sizeof(float4)));
        if (bf16) NNDEX_HIP_CHECK(hipMalloc(&cbf16, (size_t)nlist * d8 * sizeof(uint4)));
        if (!wide) NNDEX_HIP_CHECK(hipMalloc(&cimg, mfma_image_units(nlist) * sizeof(uint4)));
        uint32_t bits = 1;
        while ((1u << bits) < nlist) ++bits;
        NNDEX_HIP_CHECK(rocprim::radix_sort_pairs(sort_tmp, sort_bytes, ivf.assign, sorted_keys,
                                                  rows_iota, ivf.orig, rows, 0, bits, s));
        NNDEX_HIP_CHECK(hipMalloc(&sort_tmp, sort_bytes));

        float4 *cent = reinterpret_cast<float4 *>(ivf.centroids);
        // first centroid: the medoid row
        hipLaunchKernelGGL(colsum_kernel, dim3(cs_blocks), dim3(256), 0, s, x, rows, d4, cs_rows,
                           colpart);
        hipLaunchKernelGGL(colmean_kernel, dim3(1), dim3(256), 0, s, colpart, cs_blocks, pdims,
                           rows, mean);
        NNDEX_HIP_CHECK(hipMemsetAsync(medoid, 0, sizeof(unsigned long long), s));
        hipLaunchKernelGGL(dot_argmax_kernel, dim3(nb), dim3(256), 0, s, x, rows, d4,
                           reinterpret_cast<const float4 *>(mean), medoid);
        hipLaunchKernelGGL(copy_key_row_kernel, dim3(1), dim3(256), 0, s, x, d4, medoid, cent);
        // k-means++ seeding
        const unsigned long long seed = (unsigned long long)rows * 2654435761ull + pdims;
        NNDEX_HIP_CHECK(hipMemcpyAsync(rng, &seed, sizeof(seed), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(fill_f32_kernel, dim3(1024), dim3(256), 0, s, min_dist, (size_t)rows,
                           3.40282347e38f);
        // margin of the BF16 pre-test: delta = gamma_{d8+3} |x~||c~| + 2 xe (xm +
        // xe) for the approximate dot (8 FMA chains of d8 terms), eps =
        // gamma_{d4+2} xm^2 + u (1 + xm^2) for the reference's f32 value
        const double u = 0x1p-24, k8 = (double)d8 + 3.0, k4 = (double)d4 + 2.0;
        const double xl = x_max + x_err;
        const double kpp_margin = (k8 * u / (1.0 - k8 * u) * xl * xl + 2.0 * x_err * xl +
                                   k4 * u / (1.0 - k4 * u) * x_max * x_max +
                                   1.01 * u * (1.0 + x_max * x_max)) *
                                  (1.0 + 1e-6);
        const uint4 *kpp_xb = (xbf16 && d4 % 2 == 0 && !fp8.codes) ? xbf16 : nullptr;
        // the FP8 pre-test: gamma over eight chains of 2 d16 terms, the
        // summation tree and the scale multiply, plus |x~ - x||c| (c is a
        // row: |c| <= x_max), plus the reference value's error as above
        const double k16 = 2.0 * ((d4 + 3) / 4) + 4.0;
        const double kpp_margin8 =
            (k16 * u / (1.0 - k16 * u) * (x_max + fp8.err) * x_max + fp8.err * x_max +
             k4 * u / (1.0 - k4 * u) * x_max * x_max + 1.01 * u * (1.0 + x_max * x_max)) *
            (1.0 + 1e-6);
        // the choice of centroid c also writes its BF16 rounding (lastb) for
        // pass c + 1
        const uint32_t kb = (rows + kKppRows - 1) / kKppRows;
        if (kpp_xb)
            hipLaunchKernelGGL(bf16_round_row_kernel, dim3(1), dim3(64), 0, s, cent, d4, lastb);
        for (uint32_t c = 1; c < nlist; ++c) {
            if (fp8.codes)
                hipLaunchKernelGGL(kpp_update_fp8_kernel, dim3(kb), dim3(kKppRows), 0, s, x,
                                   fp8.codes, fp8.scale, rows, d4, cent + (size_t)(c - 1) * d4,
                                   kpp_margin8, min_dist, partial);
            else
                hipLaunchKernelGGL(kpp_update_kernel, dim3(kb), dim3(kKppRows), 0, s, x, kpp_xb,
                                   rows, d4, cent + (size_t)(c - 1) * d4, lastb, kpp_margin,
                                   min_dist, partial);
            hipLaunchKernelGGL(kpp_choose_kernel, dim3(1), dim3(kKppBlock), 0, s, x, rows, d4,
                               min_dist, partial, kb, rng, cent, c, kpp_xb ? lastb : nullptr);
        }
        NNDEX_HIP_CHECK(hipGetLastError());

        // Lloyd iterations: FP32 MFMA assignment up to 512 padded dims, BF16
        // MFMA above (centroids rounded to BF16 each iteration); either way
        // rows closer than the margin to a tie are re-checked exactly.
        // Centroids are normalised means or rows: |c| <= max(1, x_max).
        const double cmax = (x_max > 1.0 ? x_max : 1.0) * (1.0 + 1e-6);
        launch_zero_u32(reinterpret_cast<uint32_t *>(cerr), 4, s);
        if (!bf16)
            hipLaunchKernelGGL(assign_tol_kernel, dim3(1), dim3(1), 0, s, cerr, 0.0, x_max, cmax,
                               pdims, 0, tol);
        hipLaunchKernelGGL(iota_kernel, dim3(1024), dim3(256), 0, s, rows_iota, rows);
        const size_t hist_lds = (size_t)nlist * sizeof(uint32_t);
        // centroid update: a thread per float4 chunk (up to 1024), the means
        // of the chunks in LDS
        const uint32_t ublock = d4 < 1024 ? (d4 + 63) / 64 * 64 : 1024;
        const size_t upd_lds = (size_t)4 * d4 * sizeof(double);
        if (upd_lds > 64 * 1024) {
            static std::once_flag once;
            static hipError_t err = hipSuccess;
            std::call_once(once, [] {
                err = hipFuncSetAttribute((const void *)centroid_update_kernel,
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)(4 * 4096 * sizeof(double)));
            });
            NNDEX_HIP_CHECK(err);
        }
        std::vector<uint32_t> rechecked(iters, 0);
        for (uint32_t it = 0; it < iters; ++it) {
            launch_zero_u32(recheck_count, 1, s);
            if (bf16) {
                launch_zero_u32(reinterpret_cast<uint32_t *>(cerr), 2, s);
                launch_rows_to_bf16(ivf.centroids, nlist, pdims, cbf16, cerr, s);
                hipLaunchKernelGGL(assign_tol_kernel, dim3(1), dim3(1), 0, s, cerr, x_err, x_max,
                                   cmax, pdims, 1, tol);
            }
            if (wide) {
                hipLaunchKernelGGL(assign_exact_kernel, dim3((rows + 63) / 64), dim3(256), 0, s, x,
                                   rows, d4, reinterpret_cast<const float4 *>(ivf.centroids),
                                   nlist, ivf.assign);
            } else {
                MfmaScanArgs m{};
                m.x = bf16 ? xbf16 : reinterpret_cast<const uint4 *>(x);
                m.rows = rows;
                m.rows_padded = static_cast<uint32_t>(prow);
                m.dch = bf16 ? d8 : d4;
                m.q = bf16 ? cbf16 : reinterpret_cast<const uint4 *>(ivf.centroids);
                m.q_stride = m.dch;
                m.nq = nlist;
                m.mode = 2;
                // this iteration's centroids as the scan's chunk image
                {
                    uint32_t qc = 0, ru = 0;
                    mfma_chunk_shape(bf16, m.dch, 0, qc, ru, true);
                    launch_chunk_image(m.q, m.q_stride, nlist, m.dch, qc, ru, cimg, nullptr,
                                       nullptr, s);
                    m.qimg = cimg;
                }
                m.assign = ivf.assign;
                m.tol = tol;
                m.recheck = recheck;
                m.recheck_count = recheck_count;
                const uint32_t br = mfma_block_rows(bf16, m.dch, true);
                uint32_t grid = (m.rows_padded + br - 1) / br;
                if (grid > ncu) grid = ncu;
                launch_mfma_scan(m, bf16, grid, s);
                hipLaunchKernelGGL(recheck_kernel, dim3(ncu * 4), dim3(256), 0, s, x, d4,
                                   reinterpret_cast<const float4 *>(ivf.centroids), nlist, recheck,
                                   recheck_count, ivf.assign);
                NNDEX_HIP_CHECK(hipMemcpyAsync(&rechecked[it], recheck_count, sizeof(uint32_t),
                                               hipMemcpyDeviceToHost, s));
            }
            launch_zero_u32(counts, nlist, s);
            hipLaunchKernelGGL(histogram_kernel, dim3(256), dim3(256), hist_lds, s, ivf.assign,
                               rows, nlist, counts);
            hipLaunchKernelGGL(offsets_kernel, dim3(1), dim3(256), 0, s, counts, nlist,
                               ivf.offsets);
            NNDEX_HIP_CHECK(rocprim::radix_sort_pairs(sort_tmp, sort_bytes, ivf.assign,
                                                      sorted_keys, rows_iota, ivf.orig, rows, 0,
                                                      bits, s));
            hipLaunchKernelGGL(centroid_update_kernel, dim3(nlist), dim3(ublock), upd_lds, s, x,
                               d4, ivf.orig, ivf.offsets, cent);
            NNDEX_HIP_CHECK(hipGetLastError());
        }
        // reorder by the last assignment (approx.rs:64-89); orig and offsets
        // already hold the stable cluster order of that assignment
        const size_t total = prow * d4;
        uint32_t grid = (uint32_t)((total + 255) / 256);
        if (grid > 65535u * 8) grid = 65535u * 8;
        hipLaunchKernelGGL(gather_sorted_kernel, dim3(grid), dim3(256), 0, s, x, rows, d4,
                           ivf.orig, ivf.xr);
        NNDEX_HIP_CHECK(hipGetLastError());
        NNDEX_HIP_CHECK(hipStreamSynchronize(s));
        ivf.rechecked = 0;
        for (uint32_t r : rechecked) ivf.rechecked += r;
    } catch (...) {
        (void)hipStreamSynchronize(s);
        free_all();
        throw;
    }
    free_all();
    ivf.read_sizes(s);
    ivf.build_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void ivf_load(IvfIndex &ivf, const float4 *x, uint32_t rows, uint32_t pdims, uint32_t nlist,
              const float *centroids, const uint32_t *offsets, const uint32_t *orig,
              const uint32_t *assign, hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t d4 = pdims / 4;
    ivf.rows = rows;
    ivf.pdims = pdims;
    ivf.nlist = nlist;
    ivf.iters = 0;
    ivf.rechecked = 0;
    const size_t prow = padded_rows(rows);
    NNDEX_HIP_CHECK(hipMalloc(&ivf.centroids, (size_t)nlist * pdims * sizeof(float)));
    NNDEX_HIP_CHECK(hipMalloc(&ivf.offsets, (nlist + 1) * sizeof(uint32_t)));
    NNDEX_HIP_CHECK(hipMalloc(&ivf.orig, (size_t)rows * sizeof(uint32_t)));
    NNDEX_HIP_CHECK(hipMalloc(&ivf.assign, (size_t)rows * sizeof(uint32_t)));
    NNDEX_HIP_CHECK(hipMalloc(&ivf.xr, prow * d4 * sizeof(float4)));
    NNDEX_HIP_CHECK(hipMemcpyAsync(ivf.centroids, centroids, (size_t)nlist * pdims * sizeof(float),
                                   hipMemcpyHostToDevice, s));
    NNDEX_HIP_CHECK(hipMemcpyAsync(ivf.offsets, offsets, (nlist + 1) * sizeof(uint32_t),
                                   hipMemcpyHostToDevice, s));
    NNDEX_HIP_CHECK(hipMemcpyAsync(ivf.orig, orig, (size_t)rows * sizeof(uint32_t),
                                   hipMemcpyHostToDevice, s));
    NNDEX_HIP_CHECK(hipMemcpyAsync(ivf.assign, assign, (size_t)rows * sizeof(uint32_t),
                                   hipMemcpyHostToDevice, s));
    const size_t total = prow * d4;
    uint32_t grid = (uint32_t)((total + 255) / 256);
    if (grid > 65535u * 8) grid = 65535u * 8;
    hipLaunchKernelGGL(gather_sorted_kernel, dim3(grid), dim3(256), 0, s, x, rows, d4, ivf.orig,
                       ivf.xr);
    NNDEX_HIP_CHECK(hipGetLastError());
    NNDEX_HIP_CHECK(hipStreamSynchronize(s));
    ivf.read_sizes(s);
    ivf.build_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

static uint32_t ivf_bits(size_t nlist) {
    uint32_t bits = 1;
    while ((1ull << bits) < nlist) ++bits;
    return bits;
}

size_t ivf_sort_bytes(size_t pairs, size_t nlist) {
    size_t bytes = 0;
    NNDEX_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, (uint32_t *)nullptr,
                                              (uint32_t *)nullptr, (uint32_t *)nullptr,
                                              (uint32_t *)nullptr, (unsigned)pairs, 0,
                                              ivf_bits(nlist)));
    return bytes;
}

uint32_t ivf_select_k(size_t k) {
    return 2 * k - 1 <= 960 ? static_cast<uint32_t>(2 * k - 1) : static_cast<uint32_t>(k);
}

void ivf_scan_shape(size_t nq, size_t nprobe, size_t nlist, size_t rows, size_t k, uint32_t &qb,
                    uint32_t &nchunk) {
    const size_t pairs = nq * nprobe;
    // queries per wavefront: as many as the lists' pair counts use, within
    // the LDS of 4 waves x QB top-k buffers (QB 8: k <= 192, QB 4: k <= 448)
    const size_t per_list = (pairs + nlist - 1) / nlist;
    qb = per_list >= 6 ? 8 : (per_list >= 2 ? 4 : 1);
    if (qb == 8 && k > 192) qb = 4;
    if (qb == 4 && k > 448) qb = 1;
    // row chunks: about 16K wavefronts (16 per SIMD), chunks of >= 64 rows,
    // and a merge of at most 64K keys per query
    const size_t waves = (pairs + qb - 1) / qb;
    size_t nc = (16384 + waves - 1) / waves;
    const size_t avg = rows / (nlist ? nlist : 1);
    const size_t by_len = avg / 64 > 1 ? avg / 64 : 1;
    if (nc > by_len) nc = by_len;
    const size_t by_merge = 65536 / (nprobe * k) > 1 ? 65536 / (nprobe * k) : 1;
    if (nc > by_merge) nc = by_merge;
    if (nc > 64) nc = 64;
    nchunk = static_cast<uint32_t>(nc < 1 ? 1 : nc);
}

size_t ivf_scan_scratch_bytes(size_t nq, size_t nprobe, size_t nlist, size_t rows, size_t k) {
    k = ivf_select_k(k);
    uint32_t qb, nchunk;
    size_t worst = 0;
    // the part buffer is largest at the most chunks: try every batch size's
    // shape up to nq (the chunk count only falls as nq grows)
    for (size_t n = 1; n <= nq; n = n < 16 ? n + 1 : n * 2) {
        ivf_scan_shape(n, nprobe, nlist, rows, k, qb, nchunk);
        const size_t b = n * nprobe * nchunk * k * 8;
        worst = b > worst ? b : worst;
    }
    ivf_scan_shape(nq, nprobe, nlist, rows, k, qb, nchunk);
    const size_t b = nq * nprobe * nchunk * k * 8;
    worst = b > worst ? b : worst;
    const size_t pairs = nq * nprobe;
    return worst + pairs * 5 * 4 + (nlist + 1) * 4 * 4 + ivf_sort_bytes(pairs, nlist) + 1024;
}

void ivf_scan(const IvfScanArgs &a, hipStream_t s) {
    if (a.nq == 0) return;
    const uint32_t pairs = a.nq * a.nprobe;
    uint32_t grid = (pairs + 255) / 256;
    if (grid > 1024) grid = 1024;
    hipLaunchKernelGGL(pair_prep_kernel, dim3(grid), dim3(256), 0, s, a.probe, pairs, a.pair_key,
                       a.pair_val);
    hipLaunchKernelGGL(pair_base_kernel, dim3((a.nq + 255) / 256), dim3(256), 0, s, a.probe, a.nq,
                       a.nprobe, a.offsets, a.pair_base, a.totals);
    size_t bytes = a.sort_bytes;
    NNDEX_HIP_CHECK(rocprim::radix_sort_pairs(a.sort_tmp, bytes, a.pair_key, a.sorted_key,
                                              a.pair_val, a.sorted_pair, pairs, 0,
                                              ivf_bits(a.nlist), s));
    launch_zero_u32(a.pair_cnt, a.nlist, s);
    hipLaunchKernelGGL(histogram_kernel, dim3(grid < 256 ? grid : 256), dim3(256),
                       (size_t)a.nlist * sizeof(uint32_t), s, a.sorted_key, pairs, a.nlist,
                       a.pair_cnt);
    hipLaunchKernelGGL(offsets_kernel, dim3(1), dim3(256), 0, s, a.pair_cnt, a.nlist, a.pair_off);
    const uint32_t lgrid = (a.nlist + 255) / 256;
    hipLaunchKernelGGL(group_count_kernel, dim3(lgrid), dim3(256), 0, s, a.pair_cnt, a.nlist,
                       a.qb, a.grp_cnt);
    hipLaunchKernelGGL(offsets_kernel, dim3(1), dim3(256), 0, s, a.grp_cnt, a.nlist, a.grp_off);
    // upper bound on the pair groups: ceil(pairs / qb) + one partial group
    // per list that holds pairs
    const size_t gmax = (pairs + a.qb - 1) / a.qb + (a.nlist < pairs ? a.nlist : pairs);
    const size_t waves = gmax * a.nchunk;
    const uint32_t sgrid = static_cast<uint32_t>((waves + 3) / 4);
    if (a.keys_out) {
        if (a.qb == 8)
            hipLaunchKernelGGL((ivf_list_scan_kernel<8, 128, true>), dim3(sgrid), dim3(256), 0, s, a);
        else if (a.qb == 4)
            hipLaunchKernelGGL((ivf_list_scan_kernel<4, 128, true>), dim3(sgrid), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL((ivf_list_scan_kernel<1, 128, true>), dim3(sgrid), dim3(256), 0, s, a);
        NNDEX_HIP_CHECK(hipGetLastError());
        return;
    }
    const uint32_t cw = ivf_cw(a.kk);
#define NNDEX_LIST_CASE(QBV, C)                                                          \
    if (a.qb == QBV && cw == C) {                                                        \
        hipLaunchKernelGGL((ivf_list_scan_kernel<QBV, C>), dim3(sgrid), dim3(256), 0, s, a); \
        launched = true;                                                                 \
    }
    bool launched = false;
    NNDEX_LIST_CASE(8, 128)
    NNDEX_LIST_CASE(8, 256)
    NNDEX_LIST_CASE(4, 128)
    NNDEX_LIST_CASE(4, 256)
    NNDEX_LIST_CASE(4, 512)
    NNDEX_LIST_CASE(1, 128)
    NNDEX_LIST_CASE(1, 256)
    NNDEX_LIST_CASE(1, 512)
    NNDEX_LIST_CASE(1, 1024)
#undef NNDEX_LIST_CASE
    if (!launched) throw Error(NNDEX_ERR_UNSUPPORTED, "IVF search supports k <= 960");
    switch (cw) {
    case 128:
        hipLaunchKernelGGL(ivf_merge_kernel<128>, dim3(a.nq), dim3(256), 0, s, a);
        break;
    case 256:
        hipLaunchKernelGGL(ivf_merge_kernel<256>, dim3(a.nq), dim3(256), 0, s, a);
        break;
    case 512:
        hipLaunchKernelGGL(ivf_merge_kernel<512>, dim3(a.nq), dim3(256), 0, s, a);
        break;
    default:
        hipLaunchKernelGGL(ivf_merge_kernel<1024>, dim3(a.nq), dim3(256), 0, s, a);
        break;
    }
    NNDEX_HIP_CHECK(hipGetLastError());
}

// ---- large k ----------------------------------------------------------------

// Original row of arrival index `arr` of query q: the probed list holding it
// (the last probe whose base is <= arr), then the reordered -> original map.
__device__ inline uint32_t ivf_arrival_orig(const IvfScanArgs &a, uint32_t q, uint32_t arr) {
    const uint32_t *pb = a.pair_base + (size_t)q * a.nprobe;
    uint32_t lo = 0, hi = a.nprobe;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (pb[mid] <= arr)
            lo = mid;
        else
            hi = mid;
    }
    const uint32_t c = (uint32_t)a.probe[(size_t)q * a.nprobe + lo];
    return a.orig[a.offsets[c] + (arr - pb[lo])];
}

// One block per query, over its sorted top min(2k - 1, n) keys (score,
// arrival): the rows the reference heap keeps (as ivf_merge_kernel: G, then
// the tied rows up to t*, the e lowest original rows of the window evicted),
// re-keyed by original row into cand[q], their count into cnt[q].
__global__ __launch_bounds__(1024) void ivf_finish_a_kernel(IvfScanArgs a, SelectBuffers b,
                                                            uint32_t k) {
    __shared__ uint32_t h[256];
    __shared__ uint32_t so[2];
    __shared__ uint32_t g_s, u_s, e_s, tie_s, pos_s;
    const uint32_t q = blockIdx.x;
    const uint64_t *L = b.sorted + (size_t)q * b.ksel;
    const uint32_t nsel = b.cnt[q] < b.ksel ? b.cnt[q] : b.ksel;
    const uint32_t m = k < nsel ? k : nsel;
    if (threadIdx.x == 0) {
        g_s = 0;
        u_s = 0;
        e_s = 0;
        tie_s = 0;
        pos_s = 0;
        if (k >= 2 && k < nsel && (uint32_t)(L[k] >> 32) == (uint32_t)(L[k - 1] >> 32)) {
            const uint32_t sh = (uint32_t)(L[k - 1] >> 32);
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
        // t*: the k-th smallest arrival among the rows >= s
        const uint32_t tstar =
            block_select_small(u_s, k, [&](uint32_t i) { return key_row(L[i]); }, h, so);
        uint32_t local = 0;
        for (uint32_t i = threadIdx.x; i < g; i += blockDim.x) local += key_row(L[i]) > tstar;
        if (local) atomicAdd(&e_s, local);
        __syncthreads();
    }
    const uint32_t e = e_s, w = m + e;
    // the e lowest original rows of the tied window L[g, w) are evicted
    uint32_t drop_max = 0;
    if (e) {
        drop_max = block_select_small(
            w - g, e, [&](uint32_t i) { return ivf_arrival_orig(a, q, key_row(L[g + i])); }, h,
            so);
    }
    uint64_t *out = b.cand + (size_t)q * b.ksel;
    for (uint32_t i = threadIdx.x; i < w; i += blockDim.x) {
        const uint32_t id = ivf_arrival_orig(a, q, key_row(L[i]));
        if (i < g || e == 0 || id > drop_max) {
            const uint32_t p = atomicAdd(&pos_s, 1u);
            if (p < b.ksel) out[p] = (L[i] & 0xFFFFFFFF00000000ull) | (0xFFFFFFFFu - id);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) b.cnt[q] = pos_s < b.ksel ? pos_s : b.ksel;  // = m
}

__global__ __launch_bounds__(256) void ivf_finish_b_kernel(SelectBuffers b, uint32_t k,
                                                           uint64_t *__restrict__ out_idx,
                                                           float *__restrict__ out_sim,
                                                           uint64_t *__restrict__ out_len) {
    const uint32_t q = blockIdx.x;
    const uint32_t m = b.cnt[q] < k ? b.cnt[q] : k;
    const uint64_t *L = b.sorted + (size_t)q * b.ksel;
    for (uint32_t i = threadIdx.x; i < k; i += 256) {
        const uint64_t key = i < m ? L[i] : 0;
        out_idx[(size_t)q * k + i] = key ? key_row(key) : 0;
        out_sim[(size_t)q * k + i] = key ? key_sim(key) : 0.f;
    }
    if (threadIdx.x == 0) out_len[q] = m;
}

void ivf_finish_large_k(const IvfScanArgs &a, const SelectBuffers &b, uint32_t k,
                        hipStream_t s) {
    if (b.nq == 0) return;
    if (!b.n) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "the IVF selection needs per-query counts");
    hipLaunchKernelGGL(ivf_finish_a_kernel, dim3(b.nq), dim3(1024), 0, s, a, b, k);
    NNDEX_HIP_CHECK(hipGetLastError());
    select_sort(b, s);
    hipLaunchKernelGGL(ivf_finish_b_kernel, dim3(b.nq), dim3(256), 0, s, b, k, a.out_idx,
                       a.out_sim, a.out_len);
    NNDEX_HIP_CHECK(hipGetLastError());
}

void IvfIndex::read_sizes(hipStream_t s) {
    std::vector<uint32_t> off(nlist + 1);
    NNDEX_HIP_CHECK(hipMemcpyAsync(off.data(), offsets, (nlist + 1) * sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, s));
    NNDEX_HIP_CHECK(hipStreamSynchronize(s));
    std::vector<uint64_t> sz(nlist);
    for (uint32_t c = 0; c < nlist; ++c) sz[c] = off[c + 1] - off[c];
    std::sort(sz.begin(), sz.end(), [](uint64_t x, uint64_t y) { return x > y; });
    size_prefix.assign(nlist + 1, 0);
    for (uint32_t c = 0; c < nlist; ++c) size_prefix[c + 1] = size_prefix[c] + sz[c];
}

}  // namespace nndex
