// Batched search on MFMA with an exact FP32 rerank and a per-query
// certificate.
//
// Replaces the batch path of the reference GPU backend (src/gpu.rs:595-785:
// one WGSL dispatch per (row chunk, query sub-batch), full Q x rows score
// readback, a single-threaded host heap per query) with four device stages:
//
//  1. sample threshold: the batch is scored (same MFMA kernel) against a
//     strided sample of the rows; the k'-th best sample score t is the
//     threshold (at least k' rows reach it);
//  2. filter scan: MFMA scores of every (row, query) pair; pairs at or above
//     the threshold are appended to the query's candidate list (no score
//     matrix is ever written);
//  3. rerank: per query, the top k'+1 candidates by approximate score; the
//     first k' are rescored exactly in the reference's FP32 dot order and the
//     best k are returned;
//  4. certificate: with A the best approximate score left out and delta a
//     bound on |approximate - exact|, the result is exact when A + delta < the
//     k-th exact score; otherwise the query is flagged for the exact scan.
//
// The scan keeps X stationary: each wave holds its rows in registers for the
// whole batch and queries stream through LDS in chunks of 16, so the index is
// read from HBM once per batch whatever the batch size.
#include "batch.hpp"

#include <mutex>
#include "common.hpp"
#include "wave_topk.hpp"

namespace nndex {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef int v8i __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// Element traits: one 16-byte unit per lane per MFMA step group.
// ---------------------------------------------------------------------------
struct F32E {
    // four 16x16x4 f32 MFMAs over the unit's four floats (k = 4 each), each
    // into its own accumulator so the four chains are independent
    static constexpr int NACC = 4;
    static __device__ inline void mma(uint4 a, uint4 b, v4f (&c)[NACC]) {
        c[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(__builtin_bit_cast(float, a.x),
                                                    __builtin_bit_cast(float, b.x), c[0], 0, 0, 0);
        c[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(__builtin_bit_cast(float, a.y),
                                                    __builtin_bit_cast(float, b.y), c[1], 0, 0, 0);
        c[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(__builtin_bit_cast(float, a.z),
                                                    __builtin_bit_cast(float, b.z), c[2], 0, 0, 0);
        c[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(__builtin_bit_cast(float, a.w),
                                                    __builtin_bit_cast(float, b.w), c[3], 0, 0, 0);
    }
    static __device__ inline float total(const v4f (&c)[NACC], int v) {
        return (c[0][v] + c[1][v]) + (c[2][v] + c[3][v]);
    }
    static constexpr bool SCALED = false, PAIR = false;
    // the skewed epilogue only pays where MFMA steps are short (16-bit and
    // 8-bit forms); four f32 MFMAs per step already hide the epilogue
    static constexpr bool SKEW = false;
    // the mid-chunk staggered barrier (below) was measured on the BF16 form
    static constexpr bool STAGGER = false;
    static __device__ inline void mma2(uint4, uint4, uint4, uint4, uint32_t, uint32_t,
                                       v4f (&)[NACC]) {}
};
struct BF16E {
    // one 16x16x32 bf16 MFMA over the unit's eight bf16 values
    static constexpr int NACC = 1;
    static __device__ inline void mma(uint4 a, uint4 b, v4f (&c)[NACC]) {
        c[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a),
                                                       __builtin_bit_cast(v8bf, b), c[0], 0, 0, 0);
    }
    static __device__ inline float total(const v4f (&c)[NACC], int v) { return c[0][v]; }
    static constexpr bool SCALED = false, PAIR = false, SKEW = true, STAGGER = true;
    static __device__ inline void mma2(uint4, uint4, uint4, uint4, uint32_t, uint32_t,
                                       v4f (&)[NACC]) {}
};
struct FP8E {
    // one 16x16x128 f8f6f4 MFMA (both operands OCP E4M3) per two 16-byte
    // steps: lane group g supplies the 32 codes of its units 4s+g and
    // 4(s+1)+g, rows and queries alike, so the K order is the same on both
    // sides.  The exponents of the row and query scales are applied by the
    // MFMA itself as E8M0 block scales (sa: the lane's row, sb: the lane's
    // query; the same for every K block); the epilogue multiplies by the
    // scales' mantissas.
    static constexpr int NACC = 1;
    // (the staggered barrier measured no gain here and costs registers)
    static constexpr bool SCALED = true, PAIR = true, SKEW = true, STAGGER = false;
    static __device__ inline void mma(uint4, uint4, v4f (&)[NACC]) {}
    static __device__ inline void mma2(uint4 a0, uint4 a1, uint4 b0, uint4 b1, uint32_t sa,
                                       uint32_t sb, v4f (&c)[NACC]) {
        const v8i A = {(int)a0.x, (int)a0.y, (int)a0.z, (int)a0.w,
                       (int)a1.x, (int)a1.y, (int)a1.z, (int)a1.w};
        const v8i B = {(int)b0.x, (int)b0.y, (int)b0.z, (int)b0.w,
                       (int)b1.x, (int)b1.y, (int)b1.z, (int)b1.w};
        c[0] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(A, B, c[0], 0, 0, 0, (int)sa, 0,
                                                                 (int)sb);
    }
    static __device__ inline float total(const v4f (&c)[NACC], int v) { return c[0][v]; }
};

// ---------------------------------------------------------------------------
// FP8-E4M3 (OCP: bias 7, max 448, no infinities) encode with round to nearest
// even, saturating at +-448, and exact decode.
// ---------------------------------------------------------------------------
__device__ inline uint32_t fp8_e4m3(float v) {
    const uint32_t sign = (__builtin_bit_cast(uint32_t, v) >> 24) & 0x80u;
    const float a = __builtin_fabsf(v);
    uint32_t code;
    if (!(a < 464.0f)) {
        code = 0x7Eu;
    } else if (a < 0x1p-6f) {
        code = (uint32_t)__builtin_rintf(a * 512.0f);  // subnormal steps of 2^-9 (8 = 2^-6)
    } else {
        const uint32_t b = __builtin_bit_cast(uint32_t, a);
        const uint32_t e = (b >> 23) - 120u;  // 127 - 7
        const uint32_t keep = (b >> 20) & 7u, rem = b & 0xFFFFFu;
        uint32_t c = (e << 3) | keep;
        if (rem > 0x80000u || (rem == 0x80000u && (keep & 1u))) c += 1;
        code = c > 0x7Eu ? 0x7Eu : c;
    }
    return sign | code;
}
__device__ inline float fp8_e4m3_value(uint32_t c) {
    const uint32_t e = (c >> 3) & 15u, m = c & 7u;
    const float v = e == 0 ? (float)m * 0x1p-9f
                           : __builtin_bit_cast(float, ((e + 120u) << 23) | (m << 20));
    return (c & 0x80u) ? -v : v;
}

// Scale of an FP8 row or query: s = max|v| / 448, so the largest element is
// the largest code, 448, exactly (the largest elements carry most of a
// score; a power-of-two scale that leaves them off a code doubles the
// largest score error on anisotropic data), kept at or above 2^-126 (a normal
// float: its exponent is the MFMA's E8M0 block scale), 0 for a zero vector.
__host__ __device__ inline float fp8_scale(float mx) {
    if (!(mx > 0.f)) return 0.f;
    const float r = mx / 448.0f;
    return r < 0x1p-126f ? 0x1p-126f : r;
}
// Bound on |result - exact| / sum |a_k b_k| of one FP8 MFMA over its 128
// products (v_mfma_scale_f32_16x16x128_f8f6f4).  Its accumulation is not
// FP32-exact: on gfx950 random and wide-range E4M3 inputs measured up to
// 1.4e-4 (2^-12.8) of the absolute sum, 18x the FP32 bound 128 u; the bound
// used is 2^-8, 28x the largest measured.  The FP32 forms (16x16x4f32) and
// BF16 (16x16x32) measured within their FP32 bounds.
constexpr double kFp8MfmaRelErr = 0x1p-8;
// A scale s = 2^e * f, f in [1, 2): the E8M0 byte of 2^e, which the MFMA
// applies as a block scale (127 for a zero scale: its codes are 0), and the
// factor f, which the scan's epilogue applies.
__device__ inline uint32_t e8m0_of(float s) {
    const uint32_t e = (__builtin_bit_cast(uint32_t, s) >> 23) & 0xFFu;
    return e ? e : 127u;
}
__device__ inline float mant_of(float s) {
    const uint32_t b = __builtin_bit_cast(uint32_t, s);
    return (b & 0x7F800000u) ? __builtin_bit_cast(float, (b & 0x007FFFFFu) | 0x3F800000u) : 1.f;
}

// ---------------------------------------------------------------------------
// Global -> LDS copies without registers (global_load_lds_dwordx4): lane L of
// the wave copies its 16 bytes at `g` to LDS byte address m + 16 L (m is
// wave-uniform, taken through M0).  Written in asm so the compiler does not
// order every later LDS read after the copy (it cannot tell the buffers
// apart): completion is tracked by the vector memory counter, and the scan
// waits for it explicitly (wait_vm) before the barrier that publishes a
// buffer.  M0 is restored afterwards; one wait state separates its write
// from the copy.
// ---------------------------------------------------------------------------
__device__ inline uint32_t lds_off(const void *p) {
    return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void *)p;
}
__device__ inline void dma16(const uint4 *g, uint32_t m) {
    uint32_t saved;
    asm volatile(
        "s_mov_b32 %0, m0\n\t"
        "s_mov_b32 m0, %1\n\t"
        "s_nop 0\n\t"
        "global_load_lds_dwordx4 %2, off\n\t"
        "s_mov_b32 m0, %0"
        : "=&s"(saved)
        : "s"(__builtin_amdgcn_readfirstlane(m)), "v"(g)
        : "memory");
}
// s_waitcnt immediate of vmcnt(0) alone (expcnt 7, lgkmcnt 15: no wait on
// those), for the compiler's own builtin, whose waits it takes into account.
constexpr int kWaitVm0 = 0x0F70;
// Wait until at most N vector memory operations of this wave are in flight
// (they complete in issue order).
template <int N>
__device__ inline void wait_vm() {
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}
// The same for a wave-uniform count; above 16 it waits for 16 (stricter).
__device__ inline void wait_vm_dyn(uint32_t n) {
    switch (n) {
    case 0: wait_vm<0>(); break;
    case 1: wait_vm<1>(); break;
    case 2: wait_vm<2>(); break;
    case 3: wait_vm<3>(); break;
    case 4: wait_vm<4>(); break;
    case 5: wait_vm<5>(); break;
    case 6: wait_vm<6>(); break;
    case 7: wait_vm<7>(); break;
    case 8: wait_vm<8>(); break;
    case 9: wait_vm<9>(); break;
    case 10: wait_vm<10>(); break;
    case 11: wait_vm<11>(); break;
    case 12: wait_vm<12>(); break;
    case 13: wait_vm<13>(); break;
    case 14: wait_vm<14>(); break;
    case 15: wait_vm<15>(); break;
    default: wait_vm<16>(); break;
    }
}

// Chunk image of a query batch for the MFMA scan's LDS copies: chunk ch
// (queries ch*QC .. ch*QC + QC - 1) is QC*RU 16-byte units laid out exactly
// as the scan's LDS buffer: unit row c (RU of them, units c >= dch zero),
// query n at slot (n & ~15) | ((n ^ c) & 15); queries past nq are zero.  The
// thresholds and scales of the padding queries (up to nch*QC) are set so they
// pass nothing: threshold floors +inf, scales 0.
__global__ __launch_bounds__(256) void chunk_image_kernel(const uint4 *__restrict__ q,
                                                          uint32_t q_stride, uint32_t nq,
                                                          uint32_t dch, uint32_t qc, uint32_t ru,
                                                          uint32_t nch, uint4 *__restrict__ img,
                                                          float *thresh_pad, float *scale_pad) {
    const size_t per = (size_t)qc * ru, total = per * nch;
    for (size_t u = (size_t)blockIdx.x * 256 + threadIdx.x; u < total;
         u += (size_t)gridDim.x * 256) {
        const uint32_t ch = (uint32_t)(u / per), r = (uint32_t)(u % per);
        const uint32_t c = r / qc, slot = r % qc;
        const uint32_t n = (slot & ~15u) | ((slot ^ c) & 15u);
        const uint32_t qi = ch * qc + n;
        img[u] = (qi < nq && c < dch) ? q[(size_t)qi * q_stride + c] : make_uint4(0, 0, 0, 0);
    }
    if (blockIdx.x == 0)
        for (uint32_t i = nq + threadIdx.x; i < nch * qc; i += 256) {
            if (thresh_pad) thresh_pad[i] = __builtin_inff();
            if (scale_pad) scale_pad[i] = 0.f;
        }
}

// Tiled FP32 -> tiled FP8 with a per-row scale; one thread per row (lanes of a
// wave read and write 64 consecutive rows of a tile: both coalesced).
__global__ __launch_bounds__(256) void to_fp8_kernel(const float4 *__restrict__ x,
                                                     uint32_t rows_padded, uint32_t d4,
                                                     uint4 *__restrict__ y,
                                                     float *__restrict__ scale,
                                                     unsigned long long *err_sq) {
    const uint32_t row = blockIdx.x * 256 + threadIdx.x;
    double ss = 0.0;
    if (row < rows_padded) {
        const float4 *b = x + (size_t)(row / 64) * d4 * 64 + row % 64;
        float mx = 0.f;
#pragma unroll 8
        for (uint32_t c = 0; c < d4; ++c) {
            const float4 v = b[(size_t)c * 64];
            mx = __builtin_fmaxf(mx, __builtin_fmaxf(__builtin_fmaxf(__builtin_fabsf(v.x),
                                                                    __builtin_fabsf(v.y)),
                                                     __builtin_fmaxf(__builtin_fabsf(v.z),
                                                                    __builtin_fabsf(v.w))));
        }
        const float sc = fp8_scale(mx);
        const uint32_t d16 = (d4 + 3) / 4;
        uint4 *o = y + (size_t)(row / 64) * d16 * 64 + row % 64;
        for (uint32_t u = 0; u < d16; ++u) {
            uint32_t w[4];
            for (int j = 0; j < 4; ++j) {
                const uint32_t c = 4 * u + j;
                const float4 v = c < d4 ? b[(size_t)c * 64] : make_float4(0.f, 0.f, 0.f, 0.f);
                const float e[4] = {v.x, v.y, v.z, v.w};
                uint32_t word = 0;
                for (int i = 0; i < 4; ++i) {
                    const uint32_t code = sc > 0.f ? fp8_e4m3(e[i] / sc) : 0u;
                    word |= code << (8 * i);
                    const double d = (double)fp8_e4m3_value(code) * sc - (double)e[i];
                    ss += d * d;
                }
                w[j] = word;
            }
            o[(size_t)u * 64] = make_uint4(w[0], w[1], w[2], w[3]);
        }
        scale[row] = sc;
    }
    unsigned long long bits = __builtin_bit_cast(unsigned long long, ss);
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long other = __shfl_xor(bits, off);
        bits = other > bits ? other : bits;
    }
    if (__lane_id() == 0) atomicMax(err_sq, bits);
}

// ---------------------------------------------------------------------------
// BF16 copy of the tiled FP32 matrix (round to nearest even), same tiling in
// 16-byte units of eight values.  Also records each row's f64 norm bound.
// ---------------------------------------------------------------------------
__device__ inline uint32_t pack_bf16x2(float lo, float hi) {
    __bf16 a = (__bf16)lo, b = (__bf16)hi;  // RNE conversion
    return (uint32_t)__builtin_bit_cast(uint16_t, a) |
           ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
}

__global__ __launch_bounds__(256) void to_bf16_kernel(const float4 *__restrict__ x, size_t units8,
                                                      uint32_t d8, uint4 *__restrict__ y) {
    // unit u of y covers float4 chunks 2c and 2c+1 of the same row
    for (size_t u = (size_t)blockIdx.x * 256 + threadIdx.x; u < units8;
         u += (size_t)gridDim.x * 256) {
        // (rounding-error norms are accumulated per row by bf16_error_kernel)
        size_t tile = u / ((size_t)d8 * 64);
        size_t rem = u % ((size_t)d8 * 64);
        uint32_t c = rem / 64, r = rem % 64;
        size_t d4 = (size_t)d8 * 2;
        float4 a = x[(tile * d4 + 2 * c) * 64 + r];
        float4 b = x[(tile * d4 + 2 * c + 1) * 64 + r];
        y[u] = make_uint4(pack_bf16x2(a.x, a.y), pack_bf16x2(a.z, a.w), pack_bf16x2(b.x, b.y),
                          pack_bf16x2(b.z, b.w));
    }
}

// Largest per-row norm of the BF16 rounding error |x~ - x| (f64 bits).
__global__ __launch_bounds__(256) void bf16_error_kernel(const float4 *__restrict__ x,
                                                         uint32_t rows, uint32_t d4,
                                                         unsigned long long *__restrict__ out) {
    uint32_t row = blockIdx.x * 256 + threadIdx.x;
    double ss = 0.0;
    if (row < rows) {
        const float4 *b = x + (size_t)(row / 64) * d4 * 64 + row % 64;
#pragma unroll 8
        for (uint32_t c = 0; c < d4; ++c) {
            float4 v = b[(size_t)c * 64];
            float e[4] = {v.x, v.y, v.z, v.w};
            for (int j = 0; j < 4; ++j) {
                double d = (double)(float)(__bf16)e[j] - (double)e[j];
                ss += d * d;
            }
        }
    }
    unsigned long long bits = __builtin_bit_cast(unsigned long long, ss);
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long other = __shfl_xor(bits, o);
        bits = other > bits ? other : bits;
    }
    if (__lane_id() == 0) atomicMax(out, bits);
}

// Largest row norm of the tiled matrix, in f64 (bounds |x| in delta).
__global__ __launch_bounds__(256) void max_norm_kernel(const float4 *__restrict__ x,
                                                       uint32_t rows, uint32_t d4,
                                                       unsigned long long *__restrict__ out) {
    uint32_t row = blockIdx.x * 256 + threadIdx.x;
    double ss = 0.0;
    if (row < rows) {
        const float4 *b = x + (size_t)(row / 64) * d4 * 64 + row % 64;
#pragma unroll 8
        for (uint32_t c = 0; c < d4; ++c) {
            float4 v = b[(size_t)c * 64];
            ss += (double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z + (double)v.w * v.w;
        }
    }
    // non-negative doubles order like their bit patterns
    unsigned long long bits = __builtin_bit_cast(unsigned long long, ss);
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long other = __shfl_xor(bits, o);
        bits = other > bits ? other : bits;
    }
    if (__lane_id() == 0) atomicMax(out, bits);
}

// ---------------------------------------------------------------------------
// Query preparation: low-precision copy (row-major, 16-byte units; FP8 with a
// per-query scale s_q = max|q| / 448) and the per-query error bound delta of
// |approximate - exact| over any row, with x~, q~ the dequantised copies
// (x~ = x, q~ = q for FP32):
//   exact side   gamma_n |x| |q|                 (four-lane FMA chain)
//   approx side  gamma_n |x~| |q~|  (FP8: kFp8MfmaRelErr + gamma_(n/128+1)
//                                     + 3u for the scale mantissas)
//   rounding     |x~| |q~ - q| + |x~ - x| |q|
// with |x| <= xmax (largest row norm), |x~ - x| <= xe (largest row rounding
// error, measured when the copy was made) and |x~| <= xmax + xe.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void prep_queries_kernel(const float *__restrict__ q,
                                                           uint32_t nq, uint32_t pdims,
                                                           int elem, uint4 *__restrict__ qlow,
                                                           float *__restrict__ q_scale,
                                                           const unsigned long long *xmax_sq,
                                                           const unsigned long long *xerr_sq,
                                                           float *__restrict__ delta) {
    const uint32_t qi = blockIdx.x;
    if (qi >= nq) return;
    const float *src = q + (size_t)qi * pdims;
    __shared__ double red[3][256];
    __shared__ float smax[256];
    float sq = 1.f;  // FP8 query scale
    if (elem == kElemFP8) {
        float mx = 0.f;
        for (uint32_t i = threadIdx.x; i < pdims; i += 256) mx = __builtin_fmaxf(mx, __builtin_fabsf(src[i]));
        smax[threadIdx.x] = mx;
        __syncthreads();
        for (uint32_t st = 128; st > 0; st >>= 1) {
            if (threadIdx.x < st)
                smax[threadIdx.x] = __builtin_fmaxf(smax[threadIdx.x], smax[threadIdx.x + st]);
            __syncthreads();
        }
        sq = fp8_scale(smax[0]);
        if (threadIdx.x == 0) q_scale[qi] = sq;
    }
    double nq2 = 0, nl2 = 0, nd2 = 0;
    if (elem == kElemFP8) {
        const uint32_t d16 = (pdims + 15) / 16;
        for (uint32_t u = threadIdx.x; u < d16; u += 256) {
            uint32_t w[4] = {0, 0, 0, 0};
            for (int j = 0; j < 16; ++j) {
                const uint32_t i = u * 16 + j;
                const float v = i < pdims ? src[i] : 0.f;
                const uint32_t code = sq > 0.f ? fp8_e4m3(v / sq) : 0u;
                w[j / 4] |= code << (8 * (j % 4));
                const double lv = (double)fp8_e4m3_value(code) * sq;
                nq2 += (double)v * v;
                nl2 += lv * lv;
                nd2 += (lv - v) * (lv - v);
            }
            qlow[(size_t)qi * d16 + u] = make_uint4(w[0], w[1], w[2], w[3]);
        }
    } else {
        for (uint32_t i = threadIdx.x; i < pdims; i += 256) {
            const double v = src[i];
            const double lv = elem == kElemBF16 ? (double)(float)(__bf16)src[i] : v;
            nq2 += v * v;
            nl2 += lv * lv;
            nd2 += (lv - v) * (lv - v);
        }
        if (elem == kElemBF16) {
            const uint32_t d8 = pdims / 8;
            for (uint32_t c = threadIdx.x; c < d8; c += 256) {
                const float *s = src + c * 8;
                qlow[(size_t)qi * d8 + c] =
                    make_uint4(pack_bf16x2(s[0], s[1]), pack_bf16x2(s[2], s[3]),
                               pack_bf16x2(s[4], s[5]), pack_bf16x2(s[6], s[7]));
            }
        }
    }
    red[0][threadIdx.x] = nq2;
    red[1][threadIdx.x] = nl2;
    red[2][threadIdx.x] = nd2;
    __syncthreads();
    for (uint32_t st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int j = 0; j < 3; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double xm = sqrt(__builtin_bit_cast(double, *xmax_sq));
        const double u = 0x1p-24, n = pdims;
        const double gamma = n * u / (1.0 - n * u);
        const double qn = sqrt(red[0][0]), ql = sqrt(red[1][0]), qd = sqrt(red[2][0]);
        double xe = 0.0;
        if (elem == kElemBF16) {
            // the BF16 rounding error is also at most 2^-8 |x|
            xe = sqrt(__builtin_bit_cast(double, *xerr_sq));
            if (xe > 0x1p-8 * xm) xe = 0x1p-8 * xm;
        } else if (elem == kElemFP8) {
            xe = sqrt(__builtin_bit_cast(double, *xerr_sq));
        }
        const double xl = xm + xe;
        // the FP8 MFMA does not sum its 128 products within FP32's bound:
        // its error is bounded by kFp8MfmaRelErr of their absolute sum per
        // instruction, then FP32 accumulation across the n/128 instructions
        const double nc = ceil(n / 128.0) + 1.0;
        // (+3u: the epilogue's two multiplies by the scales' mantissas)
        const double ga =
            elem == kElemFP8 ? kFp8MfmaRelErr + nc * u / (1.0 - nc * u) + 3.0 * u : gamma;
        const double d = gamma * xm * qn + ga * xl * ql + xl * qd + xe * qn;
        delta[qi] = (float)(d * (1.0 + 1e-6)) + 1e-30f;
    }
}

// ---------------------------------------------------------------------------
// X-stationary MFMA scan.  NW waves per block, RW rows per wave (multiple of
// 16), MAXS 16-byte units per lane per row (units per row <= 4 * MAXS; lane
// group g = lane/16 owns units 4s+g), NT 16-query MFMA tiles per query chunk.
// Mode 0 appends candidates, mode 1 writes scores, mode 3 writes the largest
// score of each group of RW / 4 rows (a lane's rows of the wave's tiles: group
// (tile NW + wave) 4 + lane group); the AM variant (mode 2) keeps each row's
// two best queries (centroids) and writes the row argmax.
// Query chunks come from the batch's chunk image (chunk_image_kernel) by LDS
// copies issued NB - 1 chunks ahead into a ring of NB buffers, so a chunk's
// copies have NB - 2 chunks of MFMA work to land in and no registers hold
// them.
// ---------------------------------------------------------------------------
// Entries of a wave's LDS candidate queue (12 bytes each).
constexpr uint32_t kCandQueue = 128;

// LDS of one query buffer in 16-byte units: [4*steps unit rows][QC slots] of
// query units, then QC/4 units of the chunk's float threshold floors
// (candidate mode) and QC/4 units of query scales for the scaled (FP8) form.
template <typename E, int MAXS, int NT, bool AM>
constexpr uint32_t mfma_buf_units() {
    return (uint32_t)(MAXS * 4 * 16 * NT + 4 * NT + (E::SCALED ? 4 * NT : 0));
}
// Buffers in the ring: as many as fit 160 KiB beside the candidate queues
// (at most 4; the row argmax has no queues).
template <typename E, int NW, int MAXS, int NT, bool AM>
constexpr uint32_t mfma_nbuf() {
    constexpr size_t queue = AM ? 0 : (size_t)NW * kCandQueue * 12;
    constexpr size_t b = (size_t)mfma_buf_units<E, MAXS, NT, AM>() * 16, lds = 160 * 1024;
    return 4 * b + queue <= lds ? 4 : (3 * b + queue <= lds ? 3 : 2);
}
template <typename E, int NW, int MAXS, int NT, bool AM>
constexpr size_t mfma_lds_bytes() {
    return ((size_t)mfma_nbuf<E, NW, MAXS, NT, AM>() * mfma_buf_units<E, MAXS, NT, AM>() +
            (AM ? 0 : NW * kCandQueue * 12 / 16)) *
           sizeof(uint4);
}

template <typename E, int NW, int RW, int MAXS, int NT, bool AM, bool SPLIT, bool CAND>
__global__ __launch_bounds__(NW * 64, 2) void mfma_scan_kernel(MfmaScanArgs a) {
    // CAND: the candidate scan (mode 0); otherwise modes 1 and 3 (scores,
    // group maxima) or the row argmax (AM).  Separate instances keep each
    // form's registers free of the other's epilogue.
    static_assert(!(AM && CAND), "the row argmax appends no candidates");
    static_assert(!(AM && E::SCALED), "the row argmax takes unscaled scores");
    constexpr int MT = RW / 16;          // 16-row MFMA tiles per wave
    constexpr uint32_t QC = 16 * NT;     // queries per chunk
    // LDS slots per unit row: unpadded, so the 16 lanes of a group read one
    // aligned 256-byte row; query n of unit row u sits at slot
    // (n & ~15) | ((n ^ u) & 15), which keeps those reads in the row and
    // spreads the transposed commit writes over the banks
    constexpr uint32_t kStride = QC;
    constexpr uint32_t NTH = NW * 64;
    constexpr uint32_t TU = QC / 4;      // threshold units (four float floors each) per buffer
    constexpr uint32_t SU = E::SCALED ? QC / 4 : 0;  // query-scale units per buffer
    constexpr uint32_t NB = mfma_nbuf<E, NW, MAXS, NT, AM>();
    extern __shared__ uint4 qlds[];      // NB buffers x ([4*steps units][kStride] + TU + SU)
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = __lane_id();
    const uint32_t wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t g = lane >> 4, l16 = lane & 15;
    const uint32_t dch = a.dch;
    constexpr uint32_t steps = MAXS;  // launcher guarantees 4 * MAXS >= dch
    constexpr uint32_t q_units = steps * 4 * kStride;
    constexpr uint32_t buf_units = q_units + TU + SU;
    static_assert(buf_units == mfma_buf_units<E, MAXS, NT, AM>(), "LDS layout");
    const uint32_t block_rows = NW * RW;
    const uint32_t nblk = (a.rows_padded + block_rows - 1) / block_rows;
    // SPLIT: blockIdx.y takes its own share of the query chunks (a scan
    // over few rows splits the queries to fill the chip)
    const uint32_t nch_all = (a.nq + QC - 1) / QC;
    uint32_t ch0 = 0, ch1 = nch_all;
    if constexpr (SPLIT) {
        const uint32_t nch_per = (nch_all + gridDim.y - 1) / gridDim.y;
        ch0 = blockIdx.y * nch_per;
        ch1 = ch0 + nch_per < nch_all ? ch0 + nch_per : nch_all;
    }
    // Work items: the row tiles of the full rounds (every block one tile per
    // round) take all chunks; the tiles of a last, partial round are split
    // into tail_split pieces of the chunks, so the blocks finish together
    // (at 1M rows, 3,907 tiles over 256 blocks leave 67 for a 16th round)
    const uint32_t tsplit = (!SPLIT && !AM && a.tail_split > 1) ? a.tail_split : 1u;
    const uint32_t nfull = tsplit > 1 ? nblk / gridDim.x * gridDim.x : nblk;
    const uint32_t nwork = nfull + (nblk - nfull) * tsplit;
    const uint32_t tail_per = (nch_all + tsplit - 1) / tsplit;
    if (a.nq == 0 || dch == 0 || ch0 >= ch1) return;
    auto thfl = [&](uint32_t buf) {
        return reinterpret_cast<const float *>(qlds + buf * buf_units + q_units);
    };
    auto qsl = [&](uint32_t buf) {
        return reinterpret_cast<const float *>(qlds + buf * buf_units + q_units + TU);
    };
    // Mode 0 candidate queue of the wave in LDS (after the buffers): keys
    // and query ids of the pairs that passed, flushed to the per-query lists
    // when full and at the end of each row tile, so the epilogue never waits
    // on a returning atomic.
    uint64_t *cq_key = reinterpret_cast<uint64_t *>(qlds + NB * buf_units) + wid * kCandQueue;
    uint32_t *cq_qi = reinterpret_cast<uint32_t *>(qlds + NB * buf_units + NW * kCandQueue / 2) +
                      wid * kCandQueue;
    uint32_t cq_n = 0;  // wave-uniform
    auto flush = [&]() {
        wave_lds_sync();
        for (uint32_t i = lane; i < cq_n; i += 64) {
            const uint32_t qi = cq_qi[i];
            const uint64_t key = cq_key[i];
            const uint32_t pos = atomicAdd(a.cand_count + qi, 1u);
            if (pos < a.capq) a.cand[(size_t)qi * a.capq + pos] = key;
        }
        wave_lds_sync();
        cq_n = 0;
    };

    // LDS copies of one chunk from the image: NI copies of 1 KiB (64 units)
    // for the queries, one of QC float thresholds (mode 0) and one of QC
    // scales (scaled form), copy i issued by wave i % NW; my_cnt = this
    // wave's.
    constexpr uint32_t NI = (uint32_t)(NT * MAXS);
    constexpr bool stage_th = CAND;
    const uint32_t my_cnt = (NI > wid ? (NI - wid + NW - 1) / NW : 0) +
                            ((stage_th && wid == NI % NW) ? 1u : 0u) +
                            ((E::SCALED && wid == (NI + 1) % NW) ? 1u : 0u);
    const uint32_t lds0 = __builtin_amdgcn_readfirstlane(lds_off(qlds));
    auto dma_chunk = [&](uint32_t ch, uint32_t buf) {
        const uint32_t dst = lds0 + buf * buf_units * 16;
        const uint4 *src = a.qimg + (size_t)ch * q_units + lane;
        for (uint32_t i = wid; i < NI; i += NW) dma16(src + i * 64, dst + i * 64 * 16);
        // TU >= 4 and SU >= 4 lanes take part: never an empty copy
        if (stage_th && wid == NI % NW && lane < TU)
            dma16(reinterpret_cast<const uint4 *>(a.thresh_f + (size_t)ch * QC) + lane,
                  dst + q_units * 16);
        if constexpr (E::SCALED) {
            if (wid == (NI + 1) % NW && lane < SU)
                dma16(reinterpret_cast<const uint4 *>(a.q_scale + (size_t)ch * QC) + lane,
                      dst + (q_units + TU) * 16);
        }
    };

    // Mode 0 epilogue of one chunk: append the (row, query) pairs whose keys
    // reach the query's threshold to the wave's queue.  Scores are finite
    // (finite inputs).  A key at or above the threshold has a score at or
    // above the threshold's floor, so each 16-query tile first tests the
    // largest score of every lane and is skipped when none is that high.
    auto epi0 = [&](const v4f (&ac)[MT][NT][E::NACC], const float (&thv)[NT],
                    const float (&fqv)[NT], const float (&fr)[MT][4], uint32_t ch, uint32_t bt) {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const uint32_t qi = ch * QC + t * 16 + l16;
            // the threshold key's floor: every threshold key holds a score
            // in its upper half and zero below, so a score passes exactly
            // when it is at or above the floor (-0 and +0 aside, which only
            // adds a candidate); +inf for the padding queries
            const float thf = thv[t];
            // FP8: score = (acc * row mantissa) * query mantissa, so the
            // largest of acc * row mantissa times the query's decides the tile
            float sc[MT][4];
            float mx = -__builtin_inff();
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    sc[m][v] = E::total(ac[m][t], v);
                    if constexpr (E::SCALED) sc[m][v] = sc[m][v] * fr[m][v];
                    mx = __builtin_fmaxf(mx, sc[m][v]);
                }
            if constexpr (E::SCALED) mx = mx * fqv[t];
            if (__ballot(!(mx < thf)) == 0) continue;
            if constexpr (E::SCALED) {
#pragma unroll
                for (int m = 0; m < MT; ++m)
#pragma unroll
                    for (int v = 0; v < 4; ++v) sc[m][v] = sc[m][v] * fqv[t];
            }
            // Rare from here: the lane's (row, query) pairs at or above the
            // threshold as bits m*4+v, then appended a round at a time (one
            // pair per lane per round), which keeps this path short in code
            // and in time.  Rows past the index and queries past the batch
            // never pass.
            const uint32_t row0 = bt * block_rows + wid * RW + 4 * g;
            uint32_t bits = 0;
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int v = 0; v < 4; ++v) bits |= (sc[m][v] >= thf ? 1u : 0u) << (m * 4 + v);
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    if (row0 + m * 16 + v >= a.rows) bits &= ~(1u << (m * 4 + v));
            if (qi >= a.nq) bits = 0;
            for (;;) {
                const uint64_t hm = __ballot(bits != 0);
                if (hm == 0) break;
                const uint32_t n = __popcll(hm);
                if (cq_n + n > kCandQueue) flush();
                if (bits) {
                    const uint32_t b = __builtin_ctz(bits);
                    bits &= bits - 1;
                    float sv = sc[0][0];
#pragma unroll
                    for (int i = 1; i < MT * 4; ++i) sv = b == (uint32_t)i ? sc[i / 4][i % 4] : sv;
                    const uint32_t pos = cq_n + __popcll(hm & __lanemask_lt());
                    cq_key[pos] = make_key(sv, row0 + (b >> 2) * 16 + (b & 3));
                    cq_qi[pos] = qi;
                }
                cq_n += n;
            }
        }
    };
    // BF16 candidate and score scans with three or more ring buffers: the
    // chunk barrier sits inside the MFMA stream, at a different step group in
    // the two halves of the block (one wave of each SIMD's pair right after
    // the first group, the other after the last), so whichever wave waits at
    // it, its SIMD's other wave still has MFMAs to issue.  The barrier of
    // chunk ch follows every wave's last read of chunk ch - 1 and the landing
    // of chunk ch + 1's copies; after it a wave copies chunk ch + NB - 1 into
    // chunk ch - 1's buffer, which is neither of the two buffers read until
    // the next barrier (ch, ch + 1).
    constexpr bool MIDBAR = E::STAGGER && !AM && NB >= 3;
    constexpr int SW = E::PAIR ? 2 : 1;
    constexpr int NG = MAXS / SW;
    const bool bar_last = wid >= NW / 2;  // barrier after the last step group, else the first
    // Two-buffer rings: half of the waves (one of each SIMD's pair) run the
    // mode 0 epilogue one chunk late, so one wave's epilogue overlaps the
    // other's MFMAs.
    const bool late = !MIDBAR && E::SKEW && CAND && NW >= 8 && wid >= NW / 2;
    v4f accp[MT][NT][E::NACC];
    float thp[NT], fqp[NT];
    for (uint32_t w = blockIdx.x; w < nwork; w += gridDim.x) {
        uint32_t bt = w;
        if (w >= nfull) {  // block-uniform
            const uint32_t r = w - nfull, piece = r % tsplit;
            bt = nfull + r / tsplit;
            ch0 = piece * tail_per;
            ch1 = ch0 + tail_per < nch_all ? ch0 + tail_per : nch_all;
            if (ch0 >= ch1) continue;
        }
        // ---- stationary X: MT x MAXS units per lane (rows past the padded
        // count read the last row; their results are never written) ----
        uint4 xr[MT][MAXS];
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            uint32_t row = bt * block_rows + wid * RW + m * 16 + l16;
            row = row < a.rows_padded ? row : a.rows_padded - 1;
            const uint4 *base = a.x + (size_t)(row / 64) * dch * 64 + (row % 64);
#pragma unroll
            for (int s = 0; s < MAXS; ++s) {
                const uint32_t c = 4 * s + g;
                xr[m][s] = c < dch ? base[(size_t)c * 64] : make_uint4(0, 0, 0, 0);
            }
        }
        // row argmax variant only: the lane's best score and query, and the
        // second-best score (queries arrive in increasing order, so `>` keeps
        // the lowest query among equal scores; equal top scores differ by 0
        // and go to the exact re-check)
        float best[MT][4], second[MT][4];
        uint32_t bestq[MT][4];
        if constexpr (AM) {
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    best[m][v] = second[m][v] = -__builtin_inff();
                    bestq[m][v] = 0xFFFFFFFFu;
                }
        }
        // FP8: the E8M0 exponent of the scale of the lane's stationary row
        // in each row tile (applied by the MFMA), and the mantissas of the
        // scales of the rows the lane's results belong to (4g + v of each
        // tile; applied by the epilogue)
        uint32_t sa[MT];
        float fr[MT][4];
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            sa[m] = 127u;
            if constexpr (E::SCALED) {
                const uint32_t row = bt * block_rows + wid * RW + m * 16 + l16;
                sa[m] = e8m0_of(a.row_scale[row < a.rows_padded ? row : a.rows_padded - 1]);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                fr[m][v] = 1.f;
                if constexpr (E::SCALED) {
                    const uint32_t row = bt * block_rows + wid * RW + m * 16 + 4 * g + v;
                    fr[m][v] =
                        mant_of(a.row_scale[row < a.rows_padded ? row : a.rows_padded - 1]);
                }
            }
        }
        // the lane's result rows (16 m + 4 g + v of each tile) inside the
        // index, as bits m * 4 + v (mode 3 leaves the others out)
        uint32_t row_bits = 0;
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                if (bt * block_rows + wid * RW + m * 16 + 4 * g + v < a.rows) row_bits |= 1u << (m * 4 + v);
        // the stationary rows (and scales) have landed: a wait the compiler
        // sees, so it adds none for them inside the chunk loop (there it
        // would also wait for the copies in flight)
        __builtin_amdgcn_s_waitcnt(kWaitVm0);
        // the previous tile left no copies in flight (its last chunk waited
        // for all) and every wave is done reading its buffers
        __syncthreads();
        for (uint32_t j = 0; j + 1 < NB && ch0 + j < ch1; ++j) dma_chunk(ch0 + j, j);
        wait_vm<0>();
        __syncthreads();
        for (uint32_t ch = ch0; ch < ch1; ++ch) {
            const uint32_t jj = ch - ch0;
            const uint32_t cur = jj % NB;
            const uint4 *src = qlds + cur * buf_units;
            // the chunk's thresholds (mode 0) and FP8 query scales into
            // registers before the MFMAs: their LDS latency hides under them,
            // and a late epilogue runs after other waves may have begun
            // copying the next chunk into this buffer
            float thv[NT], fqv[NT];
            uint32_t sbq[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                thv[t] = 0.f;
                fqv[t] = 1.f;
                sbq[t] = 127u;
                if constexpr (CAND) thv[t] = thfl(cur)[t * 16 + l16];
                if constexpr (E::SCALED) {
                    const float sq = qsl(cur)[t * 16 + l16];
                    sbq[t] = e8m0_of(sq);
                    fqv[t] = mant_of(sq);
                }
            }
            if (late && ch > ch0) epi0(accp, thp, fqp, fr, ch - 1, bt);
            v4f acc[MT][NT][E::NACC];
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int t = 0; t < NT; ++t)
#pragma unroll
                    for (int j = 0; j < E::NACC; ++j) acc[m][t][j] = v4f{0.f, 0.f, 0.f, 0.f};
            // LDS reads run PD step groups ahead of the MFMAs; the scheduling
            // barriers keep the compiler from hoisting every group's reads.
            // An MMA covers SW steps (SW = 2 for the K=128 FP8 form).  Unit
            // row u = 4 S + g of step S sits at lane slot (l16 ^ u) & 15 =
            // l16 ^ (4 (S & 3) + g): four base pointers, constant offsets.
            constexpr int PD = NG > 2 ? 2 : 1;
            // mid-chunk form: chunks whose copies were issued past this one;
            // the barrier waits for chunk ch + 1's copies, while those of the
            // chunks after it (issued later, in order) may stay in flight
            const uint32_t ahead_mid = ch + NB - 2 < ch1 - 1 ? NB - 2 : ch1 - 1 - ch;
            const uint32_t wait_n = ahead_mid >= 2 ? (ahead_mid - 1) * my_cnt : 0;
            const uint4 *sb[4];
#pragma unroll
            for (int k = 0; k < 4; ++k)
                sb[k] = src + g * kStride + ((l16 ^ (4 * k + g)) & 15);
            uint4 qb[PD + 1][SW][NT];
            auto load_group = [&](int grp, uint4 (&dst)[SW][NT]) {
#pragma unroll
                for (int i = 0; i < SW; ++i)
#pragma unroll
                    for (int t = 0; t < NT; ++t) {
                        const int st = grp * SW + i;
                        dst[i][t] = sb[st & 3][4 * st * kStride + t * 16];
                    }
            };
#pragma unroll
            for (int p = 0; p < PD; ++p) load_group(p, qb[p]);
#pragma unroll
            for (int grp = 0; grp < NG; ++grp) {
                if (grp + PD < NG) load_group(grp + PD, qb[(grp + PD) % (PD + 1)]);
                const uint4 (&qc)[SW][NT] = qb[grp % (PD + 1)];
                const int st = grp * SW;
#pragma unroll
                for (int t = 0; t < NT; ++t)
#pragma unroll
                    for (int m = 0; m < MT; ++m) {
                        if constexpr (E::PAIR)
                            E::mma2(xr[m][st], xr[m][st + 1], qc[0][t], qc[1][t], sa[m], sbq[t],
                                    acc[m][t]);
                        else
                            E::mma(xr[m][st], qc[0][t], acc[m][t]);
                    }
                __builtin_amdgcn_sched_barrier(0);
                if (MIDBAR && (grp == 0 || grp == NG - 1)) {
                    // wave-uniform: one barrier per wave per chunk.  The two
                    // halves of the block reach it at different instructions;
                    // s_barrier counts one arrival per wave wherever the wave
                    // executes it, so every chunk's barrier still joins all
                    // NW waves (every wave runs the same chunk count)
                    if (bar_last == (grp == NG - 1)) {
                        wait_vm_dyn(wait_n);
                        __syncthreads();
                        if (ch + NB - 1 < ch1) dma_chunk(ch + NB - 1, (jj + NB - 1) % NB);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
            }
            if constexpr (CAND) {
                if (late) {
#pragma unroll
                    for (int t = 0; t < NT; ++t) {
                        thp[t] = thv[t];
                        fqp[t] = fqv[t];
#pragma unroll
                        for (int m = 0; m < MT; ++m)
#pragma unroll
                            for (int j = 0; j < E::NACC; ++j) accp[m][t][j] = acc[m][t][j];
                    }
                } else {
                    epi0(acc, thv, fqv, fr, ch, bt);
                }
            }
            // ---- mode 1 and row argmax epilogue: lane holds rows 4*g+v of
            // each tile, query 16t+l16 ----
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                if constexpr (CAND) break;
                const uint32_t qi = ch * QC + t * 16 + l16;
                if (!AM && a.mode == 3) {
                    // group maxima: the lane's 4 MT rows (16 m + 4 g + v of
                    // the wave's tiles) form group (bt NW + wid) 4 + g; rows
                    // past the index are left out (a group of padding rows
                    // gives -inf)
                    float mx = -__builtin_inff();
#pragma unroll
                    for (int m = 0; m < MT; ++m)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            float sv = E::total(acc[m][t], v);
                            if constexpr (E::SCALED) sv = (sv * fr[m][v]) * fqv[t];
                            if ((row_bits >> (m * 4 + v)) & 1u) mx = __builtin_fmaxf(mx, sv);
                        }
                    if (qi < a.nq) a.out[(size_t)qi * a.ldo + (bt * NW + wid) * 4 + g] = mx;
                    continue;
                }
#pragma unroll
                for (int m = 0; m < MT; ++m) {
                    const uint32_t row0 = bt * block_rows + wid * RW + m * 16 + 4 * g;
                    float sc[4];
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        sc[v] = E::total(acc[m][t], v);
                        if constexpr (E::SCALED) sc[v] = (sc[v] * fr[m][v]) * fqv[t];
                    }
                    if constexpr (AM) {
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const float sv = qi < a.nq ? sc[v] : -__builtin_inff();
                            const bool t = sv > best[m][v];
                            second[m][v] = t ? best[m][v] : __builtin_fmaxf(second[m][v], sv);
                            bestq[m][v] = t ? qi : bestq[m][v];
                            best[m][v] = t ? sv : best[m][v];
                        }
                    } else if (qi < a.nq) {
                        // the lane's four rows are consecutive: one 16-byte
                        // store (ldo is a multiple of 4)
                        float *dst = a.out + (size_t)qi * a.ldo + row0;
                        if (row0 + 3 < a.rows)
                            *reinterpret_cast<float4 *>(dst) = make_float4(sc[0], sc[1], sc[2], sc[3]);
                        else
                            for (int v = 0; v < 4; ++v)
                                if (row0 + v < a.rows) dst[v] = sc[v];
                    }
                }
            }
            if constexpr (!MIDBAR) {
                // two buffers: chunk ch + 1 into the buffer chunk ch - 1 used
                // (every wave passed the barrier after reading it); its copies
                // must land before the barrier at the end of this chunk
                if (ch + NB - 1 < ch1) dma_chunk(ch + NB - 1, (jj + NB - 1) % NB);
                const uint32_t ahead = ch + NB - 1 < ch1 ? NB - 1 : ch1 - 1 - ch;
                wait_vm_dyn(ahead >= 2 ? (ahead - 1) * my_cnt : 0);
                __syncthreads();
            }
        }
        if constexpr (!AM) {
            if (late) epi0(accp, thp, fqp, fr, ch1 - 1, bt);
            if (cq_n) flush();
        }
        if constexpr (AM) {
            // top two over the 16 lanes that hold the same rows (one query
            // each); a pair closer than tol goes to the exact re-check
            const float tol = *a.tol;
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    float bs = best[m][v], ss = second[m][v];
                    uint32_t bq = bestq[m][v];
                    for (int off = 1; off < 16; off <<= 1) {
                        const float obs = __shfl_xor(bs, off), oss = __shfl_xor(ss, off);
                        const uint32_t obq = __shfl_xor(bq, off);
                        const bool take = obs > bs || (obs == bs && obq < bq);
                        const float lo = take ? bs : obs;  // the loser's best
                        ss = __builtin_fmaxf(lo, __builtin_fmaxf(ss, oss));
                        bq = take ? obq : bq;
                        bs = take ? obs : bs;
                    }
                    const uint32_t row = bt * block_rows + wid * RW + m * 16 + 4 * g + v;
                    if (l16 == 0 && row < a.rows) {
                        a.assign[row] = bq;
                        if (ss > -__builtin_inff() && !(bs - ss > tol))
                            a.recheck[atomicAdd(a.recheck_count, 1u)] = row;
                    }
                }
        }
    }
}

// ---------------------------------------------------------------------------
// Sample threshold: per query, the kp-th best of `n` sample scores as a key
// threshold (keys >= thresh pass; the caller picks the rank kp so that a few
// times the rescored count of rows pass).  n < kp -> 0 (all pass).
// ---------------------------------------------------------------------------
template <int CW>
__global__ __launch_bounds__(256) void sample_threshold_kernel(const float *__restrict__ scores,
                                                               uint32_t n, uint32_t ld,
                                                               uint32_t kp,
                                                               const float *__restrict__ delta,
                                                               unsigned long long *thresh,
                                                               float *thresh_f) {
    __shared__ uint64_t slots[4][CW];
    __shared__ uint32_t cnt[4];
    const uint32_t q = blockIdx.x, lane = __lane_id(), wid = threadIdx.x >> 6;
    if (n < kp) {
        if (threadIdx.x == 0) {
            thresh[q] = 0;
            thresh_f[q] = key_floor_sim(0);
        }
        return;
    }
    WaveTopK tk;
    tk.init(slots[wid], CW, kp, nullptr);
    const float *s = scores + (size_t)q * ld;
    // independent batches of 8 loads per lane before they are offered
    constexpr int KB = 8;
    for (uint32_t i0 = wid * 64 * KB; i0 < n; i0 += 256 * KB) {
        float v[KB];
#pragma unroll
        for (int u = 0; u < KB; ++u) {
            const uint32_t i = i0 + u * 64 + lane;
            v[u] = i < n ? s[i] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < KB; ++u) {
            const uint32_t i = i0 + u * 64 + lane;
            tk.offer(i < n ? make_key(v[u], i) : 0, i < n);
        }
    }
    tk.compact();
    if (lane == 0) cnt[wid] = tk.count;
    __syncthreads();
    if (wid == 0) {
        for (uint32_t w = 1; w < 4; ++w)
            for (uint32_t i0 = 0; i0 < cnt[w]; i0 += 64) {
                uint32_t i = i0 + lane;
                tk.offer(i < cnt[w] ? slots[w][i] : 0, i < cnt[w]);
            }
        tk.compact();
        if (lane == 0) {
            // At least kp sample rows score >= t; rows under t are bounded by
            // t + delta in the certificate, so no lowering is needed here.
            const float lowered = key_sim(slots[0][kp - 1]);
            thresh[q] = (unsigned long long)ordered_bits(lowered) << 32;
            thresh_f[q] = key_floor_sim(thresh[q]);
        }
    }
}

// ---------------------------------------------------------------------------
// Rerank + certificate, one block per query.
// ---------------------------------------------------------------------------
template <int CW>
__global__ __launch_bounds__(256) void rerank_kernel(RerankArgs a) {
    __shared__ uint64_t slots[4][CW];
    __shared__ uint32_t cnt[4];
    __shared__ uint64_t exact[CW];
    // per wave: 64 rows x (8 + 1) float4 (the large over-fetch form only)
    __shared__ float4 stage[CW >= 512 ? 4 : 1][CW >= 512 ? 64 * 9 : 1];
    const uint32_t q = blockIdx.x, lane = __lane_id(), wid = threadIdx.x >> 6;
    const uint32_t total = a.cand_count[q];
    const bool overflow = total > a.capq;
    const uint32_t n = overflow ? a.capq : total;
    const uint32_t kp1 = a.kp + 1;
    const uint64_t *c = a.cand + (size_t)q * a.capq;

    WaveTopK tk;
    tk.init(slots[wid], CW, kp1, nullptr);
    constexpr int KB = 8;  // independent batches of 8 loads per lane
    for (uint32_t i0 = wid * 64 * KB; i0 < n; i0 += 256 * KB) {
        uint64_t v[KB];
#pragma unroll
        for (int u = 0; u < KB; ++u) {
            const uint32_t i = i0 + u * 64 + lane;
            v[u] = i < n ? c[i] : 0;
        }
#pragma unroll
        for (int u = 0; u < KB; ++u) tk.offer(v[u], i0 + u * 64 + lane < n);
    }
    tk.compact();
    if (lane == 0) cnt[wid] = tk.count;
    __syncthreads();
    if (wid == 0) {
        for (uint32_t w = 1; w < 4; ++w)
            for (uint32_t i0 = 0; i0 < cnt[w]; i0 += 64) {
                uint32_t i = i0 + lane;
                tk.offer(i < cnt[w] ? slots[w][i] : 0, i < cnt[w]);
            }
        tk.compact();
        if (lane == 0) cnt[0] = tk.count;
    }
    __syncthreads();
    const uint32_t m = cnt[0];                       // approx list length (<= k'+1)
    const uint32_t nres = m < a.kp ? m : a.kp;       // rescored
    // Exact rescoring in the reference dot order, one lane per candidate.
    if constexpr (CW >= 512) {
        // Large over-fetch: each wave copies its 64 candidates' rows
        // through LDS in segments of kSeg float4 (kSeg lanes read one row's
        // 128-byte run, rather than 64 lanes each reading its own row),
        // then every lane runs its row's four FMA chains over the
        // segment.  The loads of the next two segments are in flight
        // while one is summed.
        constexpr uint32_t kSeg = 8, kPad = kSeg + 1;  // padded rows: no bank conflicts
        float4 *stg = stage[CW >= 512 ? wid : 0];
        const float4 *qrow = a.q + (size_t)q * a.d4;
        for (uint32_t i0 = 0; i0 < CW; i0 += 256) {
            const uint32_t wbase = i0 + wid * 64;
            const uint32_t i = wbase + lane;
            float l0 = 0.f, l1 = 0.f, l2 = 0.f, l3 = 0.f;
            if (wbase < nres) {
                // the rows this lane copies: candidate 8 j + lane / 8 of the wave
                uint32_t crow[8];
    #pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const uint32_t ci = wbase + 8 * j + lane / kSeg;
                    crow[j] = ci < nres ? key_row(slots[0][ci]) : 0xFFFFFFFFu;
                }
                // two register sets: segment s + 2 is loading while s is summed
                float4 va[8], vb[8];
                auto load = [&](float4(&v)[8], uint32_t c0) {
                    const uint32_t cc = c0 + lane % kSeg;
    #pragma unroll
                    for (int j = 0; j < 8; ++j)
                        v[j] = crow[j] != 0xFFFFFFFFu && cc < a.d4 ? a.x[(size_t)crow[j] * a.d4 + cc]
                                                                   : make_float4(0.f, 0.f, 0.f, 0.f);
                };
                auto segment = [&](float4(&v)[8], uint32_t c0) {
                    wave_lds_sync();  // the previous segment is summed
    #pragma unroll
                    for (int j = 0; j < 8; ++j) stg[(8 * j + lane / kSeg) * kPad + lane % kSeg] = v[j];
                    wave_lds_sync();
                    if (c0 + 2 * kSeg < a.d4) load(v, c0 + 2 * kSeg);
                    const uint32_t nc = a.d4 - c0 < kSeg ? a.d4 - c0 : kSeg;
                    for (uint32_t cc = 0; cc < nc; ++cc) {
                        const float4 xv = stg[lane * kPad + cc], qv = qrow[c0 + cc];
                        l0 = __builtin_fmaf(xv.x, qv.x, l0);
                        l1 = __builtin_fmaf(xv.y, qv.y, l1);
                        l2 = __builtin_fmaf(xv.z, qv.z, l2);
                        l3 = __builtin_fmaf(xv.w, qv.w, l3);
                    }
                };
                load(va, 0);
                if (kSeg < a.d4) load(vb, kSeg);
                for (uint32_t c0 = 0; c0 < a.d4; c0 += 2 * kSeg) {
                    segment(va, c0);
                    if (c0 + kSeg < a.d4) segment(vb, c0 + kSeg);
                }
            }
            if (i < CW) exact[i] = i < nres ? make_key((l0 + l1) + (l2 + l3), key_row(slots[0][i])) : 0;
        }
    } else {
        // a few rows per query: one thread streams each candidate row
        for (uint32_t i = threadIdx.x; i < CW; i += 256) {
            uint64_t key = 0;
            if (i < nres) {
                const uint32_t row = key_row(slots[0][i]);
                const float4 *xb = a.x + (size_t)row * a.d4;
                const float4 *qb = a.q + (size_t)q * a.d4;
                float l0 = 0.f, l1 = 0.f, l2 = 0.f, l3 = 0.f;
                // unrolled so a 128-byte run's loads are issued together
#pragma unroll 8
                for (uint32_t cc = 0; cc < a.d4; ++cc) {
                    float4 xv = xb[cc], qv = qb[cc];
                    l0 = __builtin_fmaf(xv.x, qv.x, l0);
                    l1 = __builtin_fmaf(xv.y, qv.y, l1);
                    l2 = __builtin_fmaf(xv.z, qv.z, l2);
                    l3 = __builtin_fmaf(xv.w, qv.w, l3);
                }
                key = make_key((l0 + l1) + (l2 + l3), row);
            }
            exact[i] = key;
        }
    }
    __syncthreads();
    if (wid == 0) {
        WaveTopK ex;
        ex.init(exact, CW, a.k, nullptr);
        ex.count = nres;
        ex.sort_desc();
        const bool enough = nres >= a.k;
        for (uint32_t i = lane; i < a.k; i += 64) {
            uint64_t key = i < nres ? exact[i] : 0;
            a.out_idx[(size_t)q * a.k + i] = key ? key_row(key) : 0;
            a.out_sim[(size_t)q * a.k + i] = key ? key_sim(key) : 0.f;
        }
        if (lane == 0) {
            // Rows left out of the rescoring: candidates ranked after k'
            // (approx <= slots[k']) and rows under the threshold (approx <
            // t_low).  With no left-out rows the answer is exact outright.
            const unsigned long long th = a.thresh[q];
            bool ok = !overflow && enough;
            // a tie across the k-th score: the reference's heap picks among
            // the tied rows by arrival order, which the exact scan reproduces
            if (ok && a.k > 1 && nres > a.k &&
                (uint32_t)(exact[a.k] >> 32) == (uint32_t)(exact[a.k - 1] >> 32))
                ok = false;
            const bool left_out = m > a.kp || th != 0;
            if (ok && left_out) {
                const double bound = m > a.kp ? (double)key_sim(slots[0][a.kp])
                                              : (double)from_ordered((uint32_t)(th >> 32));
                ok = bound + (double)a.delta[q] < (double)key_sim(exact[a.k - 1]);
            }
            a.flags[q] = ok ? 0 : 1;
        }
    }
}

// Second rerank for queries the first certificate left open: every candidate
// (all rows whose approximate key passed the threshold) is rescored exactly.
// The rows left out then all have approx < t_low, so exact < t_low + delta,
// and the result is exact when t_low + delta < the k-th exact score.
template <int CW>
__global__ __launch_bounds__(256) void rerank_full_kernel(RerankArgs a) {
    __shared__ uint64_t slots[4][CW];
    __shared__ uint32_t cnt[4];
    const uint32_t q = blockIdx.x, lane = __lane_id(), wid = threadIdx.x >> 6;
    if (a.flags[q] == 0) return;
    const uint32_t total = a.cand_count[q];
    if (total > a.capq) return;  // overflow: stays flagged for the exact scan
    const uint64_t *c = a.cand + (size_t)q * a.capq;
    const float4 *qb = a.q + (size_t)q * a.d4;
    WaveTopK tk;
    tk.init(slots[wid], CW, a.k + 1, nullptr);  // one past k: ties across the k-th
    for (uint32_t i0 = wid * 64; i0 < total; i0 += 256) {
        const uint32_t i = i0 + lane;
        uint64_t key = 0;
        if (i < total) {
            const uint32_t row = key_row(c[i]);
            const float4 *xb = a.x + (size_t)row * a.d4;
            float l0 = 0.f, l1 = 0.f, l2 = 0.f, l3 = 0.f;
#pragma unroll 8
            for (uint32_t cc = 0; cc < a.d4; ++cc) {
                float4 xv = xb[cc], qv = qb[cc];
                l0 = __builtin_fmaf(xv.x, qv.x, l0);
                l1 = __builtin_fmaf(xv.y, qv.y, l1);
                l2 = __builtin_fmaf(xv.z, qv.z, l2);
                l3 = __builtin_fmaf(xv.w, qv.w, l3);
            }
            key = make_key((l0 + l1) + (l2 + l3), row);
        }
        tk.offer(key, i < total);
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
    if (tk.count < a.k) return;
    const unsigned long long th = a.thresh[q];
    const double t_low = th ? (double)from_ordered((uint32_t)(th >> 32)) : -1e300;
    const bool ok = th == 0 || t_low + (double)a.delta[q] < (double)key_sim(slots[0][a.k - 1]);
    if (!ok) return;
    if (a.k > 1 && tk.count > a.k &&
        (uint32_t)(slots[0][a.k] >> 32) == (uint32_t)(slots[0][a.k - 1] >> 32))
        return;  // tie across the k-th score: left to the exact scan
    for (uint32_t i = lane; i < a.k; i += 64) {
        a.out_idx[(size_t)q * a.k + i] = key_row(slots[0][i]);
        a.out_sim[(size_t)q * a.k + i] = key_sim(slots[0][i]);
    }
    if (lane == 0) a.flags[q] = 0;
}

// Tiled -> row-major, one block per 64-row tile: 16 chunks of the tile's 64
// rows staged in LDS (1 KiB coalesced reads), then written row by row (16
// threads per 256-byte row segment).
__global__ __launch_bounds__(256) void tiled_to_rows_kernel(const float4 *__restrict__ x,
                                                            uint32_t rows, uint32_t d4,
                                                            float4 *__restrict__ y) {
    __shared__ float4 t[16][64];
    const uint32_t tile = blockIdx.x;
    const float4 *src = x + (size_t)tile * d4 * 64;
    for (uint32_t c0 = 0; c0 < d4; c0 += 16) {
        const uint32_t nc = d4 - c0 < 16 ? d4 - c0 : 16;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nc * 64; i += 256)
            t[i / 64][i % 64] = src[(size_t)(c0 + i / 64) * 64 + i % 64];
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nc * 64; i += 256) {
            const uint32_t r = i / nc, c = i % nc, row = tile * 64 + r;
            if (row < rows) y[(size_t)row * d4 + c0 + c] = t[c][r];
        }
    }
}

void launch_tiled_to_rows(const float4 *x, uint32_t rows, uint32_t d4, float4 *y, hipStream_t s) {
    const uint32_t tiles = (rows + 63) / 64;
    if (tiles == 0) return;
    hipLaunchKernelGGL(tiled_to_rows_kernel, dim3(tiles), dim3(256), 0, s, x, rows, d4, y);
    NNDEX_HIP_CHECK(hipGetLastError());
}

// Strided row sample of the tiled matrix: sample row i = row i * stride.
__global__ __launch_bounds__(256) void gather_rows_kernel(const float4 *__restrict__ x, uint32_t d4,
                                                          uint32_t stride, uint32_t rows_s,
                                                          float4 *__restrict__ dst) {
    const size_t total = (size_t)((rows_s + 63) / 64) * 64 * d4;
    for (size_t u = (size_t)blockIdx.x * 256 + threadIdx.x; u < total;
         u += (size_t)gridDim.x * 256) {
        size_t tile = u / ((size_t)d4 * 64), rem = u % ((size_t)d4 * 64);
        uint32_t c = rem / 64, r = rem % 64;
        size_t i = tile * 64 + r;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (i < rows_s) {
            size_t row = i * stride;
            v = x[((row / 64) * d4 + c) * 64 + row % 64];
        }
        dst[u] = v;
    }
}

// Row-major rows -> BF16 units; one block per row.  err_sq: largest row norm
// of the rounding error, f64 bits (non-negative doubles order as their bits).
__global__ __launch_bounds__(256) void rows_to_bf16_kernel(const float *__restrict__ rows,
                                                           uint32_t pdims,
                                                           uint4 *__restrict__ out,
                                                           unsigned long long *err_sq) {
    __shared__ double red[256];
    const uint32_t r = blockIdx.x, d8 = pdims / 8;
    const float *src = rows + (size_t)r * pdims;
    double ss = 0.0;
    for (uint32_t u = threadIdx.x; u < d8; u += 256) {
        const float *e = src + u * 8;
        out[(size_t)r * d8 + u] = make_uint4(pack_bf16x2(e[0], e[1]), pack_bf16x2(e[2], e[3]),
                                             pack_bf16x2(e[4], e[5]), pack_bf16x2(e[6], e[7]));
        for (int j = 0; j < 8; ++j) {
            const double d = (double)(float)(__bf16)e[j] - (double)e[j];
            ss += d * d;
        }
    }
    red[threadIdx.x] = ss;
    __syncthreads();
    for (uint32_t st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(err_sq, __builtin_bit_cast(unsigned long long, red[0]));
}

// Graph-safe clear (a kernel, not a memset node).
__global__ void zero_u32_kernel(uint32_t *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        p[i] = 0;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

void launch_gather_rows(const float4 *x, uint32_t d4, uint32_t stride, uint32_t rows_s,
                        float4 *dst, hipStream_t s) {
    size_t total = (size_t)((rows_s + 63) / 64) * 64 * d4;
    uint32_t grid = (uint32_t)((total + 255) / 256);
    if (grid > 65535u * 8) grid = 65535u * 8;
    hipLaunchKernelGGL(gather_rows_kernel, dim3(grid), dim3(256), 0, s, x, d4, stride, rows_s,
                       dst);
    NNDEX_HIP_CHECK(hipGetLastError());
}

void launch_zero_u32(uint32_t *p, size_t n, hipStream_t s) {
    uint32_t grid = (uint32_t)((n + 255) / 256);
    if (grid > 4096) grid = 4096;
    if (grid == 0) grid = 1;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(grid), dim3(256), 0, s, p, n);
    NNDEX_HIP_CHECK(hipGetLastError());
}

void launch_rows_to_bf16(const float *rows, uint32_t n, uint32_t pdims, uint4 *out,
                         unsigned long long *err_sq, hipStream_t s) {
    hipLaunchKernelGGL(rows_to_bf16_kernel, dim3(n), dim3(256), 0, s, rows, pdims, out, err_sq);
    NNDEX_HIP_CHECK(hipGetLastError());
}

void launch_to_bf16(const float4 *x, size_t rows_padded, uint32_t d8, uint4 *y, hipStream_t s) {
    size_t units = rows_padded * d8;
    uint32_t grid = (uint32_t)((units + 255) / 256);
    if (grid > 65535u * 8) grid = 65535u * 8;
    hipLaunchKernelGGL(to_bf16_kernel, dim3(grid), dim3(256), 0, s, x, units, d8, y);
    NNDEX_HIP_CHECK(hipGetLastError());
}

void launch_bf16_error(const float4 *x, uint32_t rows, uint32_t d4, unsigned long long *out,
                       hipStream_t s) {
    hipLaunchKernelGGL(bf16_error_kernel, dim3((rows + 255) / 256), dim3(256), 0, s, x, rows, d4,
                       out);
    NNDEX_HIP_CHECK(hipGetLastError());
}

void launch_max_norm(const float4 *x, uint32_t rows, uint32_t d4, unsigned long long *out,
                     hipStream_t s) {
    hipLaunchKernelGGL(max_norm_kernel, dim3((rows + 255) / 256), dim3(256), 0, s, x, rows, d4,
                       out);
    NNDEX_HIP_CHECK(hipGetLastError());
}

void launch_prep_queries(const float *q, uint32_t nq, uint32_t pdims, int elem, uint4 *qlow,
                         float *q_scale, const unsigned long long *xmax_sq,
                         const unsigned long long *xerr_sq, float *delta, hipStream_t s) {
    hipLaunchKernelGGL(prep_queries_kernel, dim3(nq), dim3(256), 0, s, q, nq, pdims, elem, qlow,
                       q_scale, xmax_sq, xerr_sq, delta);
    NNDEX_HIP_CHECK(hipGetLastError());
}

void launch_to_fp8(const float4 *x, size_t rows_padded, uint32_t d4, uint4 *y, float *scale,
                   unsigned long long *err_sq, hipStream_t s) {
    const uint32_t grid = (uint32_t)((rows_padded + 255) / 256);
    hipLaunchKernelGGL(to_fp8_kernel, dim3(grid), dim3(256), 0, s, x, (uint32_t)rows_padded, d4,
                       y, scale, err_sq);
    NNDEX_HIP_CHECK(hipGetLastError());
}

// One wavefront's chain of `calls` MFMAs in the scan's element form over
// caller operands (lane-major per call: FP32 one float per lane, BF16 eight
// values, FP8 32 codes with unit E8M0 scales), the 16 x 16 FP32 result
// row-major in `out` (row 4 (lane / 16) + v, column lane % 16).  Tests check
// each form's accumulation error against the bound the certificate assumes.
__global__ __launch_bounds__(64) void mfma_tile_kernel(int elem, const uint4 *__restrict__ a,
                                                       const uint4 *__restrict__ b,
                                                       uint32_t calls, float *__restrict__ out) {
    const uint32_t l = threadIdx.x;
    v4f c[4] = {v4f{0.f, 0.f, 0.f, 0.f}, v4f{0.f, 0.f, 0.f, 0.f}, v4f{0.f, 0.f, 0.f, 0.f},
                v4f{0.f, 0.f, 0.f, 0.f}};
    for (uint32_t k = 0; k < calls; ++k) {
        if (elem == kElemF32) {
            const float *af = reinterpret_cast<const float *>(a), *bf = reinterpret_cast<const float *>(b);
            c[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[k * 64 + l], bf[k * 64 + l], c[0], 0, 0, 0);
        } else if (elem == kElemBF16) {
            v4f one[1] = {c[0]};
            BF16E::mma(a[k * 64 + l], b[k * 64 + l], one);
            c[0] = one[0];
        } else {
            v4f one[1] = {c[0]};
            FP8E::mma2(a[(k * 64 + l) * 2], a[(k * 64 + l) * 2 + 1], b[(k * 64 + l) * 2],
                       b[(k * 64 + l) * 2 + 1], 127u, 127u, one);
            c[0] = one[0];
        }
    }
    for (int v = 0; v < 4; ++v) out[(4 * (l / 16) + v) * 16 + l % 16] = c[0][v];
}

void launch_mfma_tile(int elem, const void *a, const void *b, uint32_t calls, float *out,
                      hipStream_t s) {
    if (elem < kElemF32 || elem > kElemFP8) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "element form");
    hipLaunchKernelGGL(mfma_tile_kernel, dim3(1), dim3(64), 0, s, elem,
                       static_cast<const uint4 *>(a), static_cast<const uint4 *>(b), calls, out);
    NNDEX_HIP_CHECK(hipGetLastError());
}

template <typename E, int NW, int RW, int MAXS, int NT, bool AM, bool SPLIT, bool CAND>
static void launch_mfma_split(const MfmaScanArgs &a, uint32_t grid, hipStream_t s) {
    // the ring of query buffers, then the waves' candidate queues
    constexpr size_t lds = mfma_lds_bytes<E, NW, MAXS, NT, AM>();
    static_assert(lds <= 160 * 1024, "LDS of the MFMA scan");
    if (!a.qimg) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "the MFMA scan needs a chunk image");
    auto fn = mfma_scan_kernel<E, NW, RW, MAXS, NT, AM, SPLIT, CAND>;
    // once per instantiation (also keeps the call out of graph captures);
    // serving executors launch from several threads, so the first call sets
    // it and every other waits until it is set
    static std::once_flag attr_once;
    static hipError_t attr_err = hipSuccess;
    std::call_once(attr_once, [&] {
        attr_err = hipFuncSetAttribute((const void *)fn,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    });
    NNDEX_HIP_CHECK(attr_err);
    hipLaunchKernelGGL(fn, dim3(grid, SPLIT ? a.qsplit : 1), dim3(NW * 64), lds, s, a);
    NNDEX_HIP_CHECK(hipGetLastError());
}

template <typename E, int NW, int RW, int MAXS, int NT, bool AM = false>
static void launch_mfma_t(const MfmaScanArgs &a, uint32_t grid, hipStream_t s) {
    if constexpr (!AM) {
        if (a.mode != 0 && a.mode != 1 && a.mode != 3)
            throw Error(NNDEX_ERR_INVALID_ARGUMENT, "MFMA scan mode");
        if (a.mode == 0) {
            if (a.qsplit > 1) return launch_mfma_split<E, NW, RW, MAXS, NT, AM, true, true>(a, grid, s);
            return launch_mfma_split<E, NW, RW, MAXS, NT, AM, false, true>(a, grid, s);
        }
        if (a.qsplit > 1) return launch_mfma_split<E, NW, RW, MAXS, NT, AM, true, false>(a, grid, s);
    }
    launch_mfma_split<E, NW, RW, MAXS, NT, AM, false, false>(a, grid, s);
}

// Tile shapes per step count (a row of `dch` 16-byte units runs the smallest
// step count with 4 * steps >= dch; extra units are zero): the 16-query MFMA
// tiles per chunk (nt) and rows per wave (rw) of each element type for the
// candidate and score scans, with an alternative shape for the autotuner
// (variant 1; equal to the default where no other shape fits 256 VGPRs
// without heavy spills), and the row argmax's own shapes (its epilogue keeps
// two keys per row).  Defaults chosen by timing at 1M x 128-1024 with 2048
// queries.
struct MfmaTiles {
    uint32_t steps;
    int f32_nt, bf_rw, bf_nt, f8_rw, f8_nt;  // mode 0 / 1, variant 0
    int f32_nt1, bf_rw1, bf_nt1, f8_rw1, f8_nt1;  // variant 1
    int am_f32_nt, am_bf_rw, am_bf_nt;       // mode 2
};
static constexpr MfmaTiles kMfmaTiles[] = {
    {2, 4, 32, 4, 32, 4, 2, 32, 2, 32, 2, 4, 32, 4},
    {4, 4, 32, 4, 32, 2, 2, 32, 2, 16, 2, 4, 32, 4},
    {6, 2, 32, 2, 32, 2, 4, 32, 4, 16, 2, 2, 32, 2},
    {8, 2, 32, 2, 16, 2, 4, 32, 4, 32, 2, 2, 32, 2},
    {12, 2, 48, 2, 32, 2, 1, 32, 4, 16, 2, 2, 16, 4},
    {16, 2, 32, 1, 32, 1, 1, 16, 2, 16, 2, 2, 16, 2},
    {24, 2, 32, 1, 16, 1, 1, 16, 2, 16, 1, 1, 16, 1},
    {32, 1, 16, 1, 16, 1, 1, 16, 1, 16, 1, 1, 16, 1},
};
constexpr int kMfmaTileCount = sizeof(kMfmaTiles) / sizeof(kMfmaTiles[0]);

static int mfma_tile_index(uint32_t dch) {
    for (int i = 0; i < kMfmaTileCount; ++i)
        if (4 * kMfmaTiles[i].steps >= dch) return i;
    return -1;
}

uint32_t mfma_variants(int elem, uint32_t dch) {
    const int i = mfma_tile_index(dch);
    if (i < 0) return 1;
    const MfmaTiles &t = kMfmaTiles[i];
    if (elem == kElemFP8) return t.f8_rw1 != t.f8_rw || t.f8_nt1 != t.f8_nt ? 2 : 1;
    if (elem == kElemBF16) return t.bf_rw1 != t.bf_rw || t.bf_nt1 != t.bf_nt ? 2 : 1;
    return t.f32_nt1 != t.f32_nt ? 2 : 1;
}

uint32_t mfma_block_rows(int elem, uint32_t dch, bool argmax, int variant) {
    const int i = mfma_tile_index(dch);
    if (i < 0) return 8 * 16;
    const MfmaTiles &t = kMfmaTiles[i];
    if (elem == kElemFP8) return 8 * (variant ? t.f8_rw1 : t.f8_rw);
    if (elem == kElemBF16) return 8 * (argmax ? t.am_bf_rw : (variant ? t.bf_rw1 : t.bf_rw));
    return 8 * 16;
}

template <int I>
static void launch_mfma_case(const MfmaScanArgs &a, int elem, uint32_t grid, hipStream_t s,
                             int variant) {
    constexpr MfmaTiles t = kMfmaTiles[I];
    constexpr int ST = (int)t.steps;
    if (a.mode == 2) {
        if (elem == kElemF32) return launch_mfma_t<F32E, 8, 16, ST, t.am_f32_nt, true>(a, grid, s);
        if (elem == kElemBF16)
            return launch_mfma_t<BF16E, 8, t.am_bf_rw, ST, t.am_bf_nt, true>(a, grid, s);
        throw Error(NNDEX_ERR_UNSUPPORTED, "no FP8 row argmax scan");
    }
    if (elem == kElemFP8) {
        if (variant) return launch_mfma_t<FP8E, 8, t.f8_rw1, ST, t.f8_nt1>(a, grid, s);
        return launch_mfma_t<FP8E, 8, t.f8_rw, ST, t.f8_nt>(a, grid, s);
    }
    if (elem == kElemBF16) {
        if (variant) return launch_mfma_t<BF16E, 8, t.bf_rw1, ST, t.bf_nt1>(a, grid, s);
        return launch_mfma_t<BF16E, 8, t.bf_rw, ST, t.bf_nt>(a, grid, s);
    }
    if (variant) return launch_mfma_t<F32E, 8, 16, ST, t.f32_nt1>(a, grid, s);
    return launch_mfma_t<F32E, 8, 16, ST, t.f32_nt>(a, grid, s);
}

void mfma_chunk_shape(int elem, uint32_t dch, int variant, uint32_t &qc, uint32_t &ru,
                      bool argmax) {
    if (variant < 0 || variant > 1 || (uint32_t)variant >= mfma_variants(elem, dch)) variant = 0;
    const int i = mfma_tile_index(dch);
    if (i < 0)
        throw Error(NNDEX_ERR_UNSUPPORTED, "padded dims above 512 (f32) / 1024 (bf16) / 2048 "
                                           "(fp8) are not supported by the MFMA scan");
    const MfmaTiles &t = kMfmaTiles[i];
    int nt;
    if (argmax)
        nt = elem == kElemBF16 ? t.am_bf_nt : t.am_f32_nt;
    else if (elem == kElemFP8)
        nt = variant ? t.f8_nt1 : t.f8_nt;
    else if (elem == kElemBF16)
        nt = variant ? t.bf_nt1 : t.bf_nt;
    else
        nt = variant ? t.f32_nt1 : t.f32_nt;
    qc = 16u * (uint32_t)nt;
    ru = 4u * t.steps;
}

void launch_chunk_image(const uint4 *q, uint32_t q_stride, uint32_t nq, uint32_t dch, uint32_t qc,
                        uint32_t ru, uint4 *img, float *thresh_pad, float *scale_pad,
                        hipStream_t s) {
    if (nq == 0) return;
    const uint32_t nch = (nq + qc - 1) / qc;
    const size_t total = (size_t)nch * qc * ru;
    if (total > mfma_image_units(nq))
        throw Error(NNDEX_ERR_INVALID_ARGUMENT, "chunk image larger than its buffer");
    uint32_t grid = (uint32_t)((total + 255) / 256);
    if (grid > 2048) grid = 2048;
    hipLaunchKernelGGL(chunk_image_kernel, dim3(grid), dim3(256), 0, s, q, q_stride, nq, dch, qc,
                       ru, nch, img, thresh_pad, scale_pad);
    NNDEX_HIP_CHECK(hipGetLastError());
}

void launch_mfma_scan(const MfmaScanArgs &a, int elem, uint32_t grid, hipStream_t s,
                      int variant) {
    if (variant < 0 || variant > 1 || (uint32_t)variant >= mfma_variants(elem, a.dch))
        variant = 0;
    switch (mfma_tile_index(a.dch)) {
    case 0: return launch_mfma_case<0>(a, elem, grid, s, variant);
    case 1: return launch_mfma_case<1>(a, elem, grid, s, variant);
    case 2: return launch_mfma_case<2>(a, elem, grid, s, variant);
    case 3: return launch_mfma_case<3>(a, elem, grid, s, variant);
    case 4: return launch_mfma_case<4>(a, elem, grid, s, variant);
    case 5: return launch_mfma_case<5>(a, elem, grid, s, variant);
    case 6: return launch_mfma_case<6>(a, elem, grid, s, variant);
    case 7: return launch_mfma_case<7>(a, elem, grid, s, variant);
    default:
        throw Error(NNDEX_ERR_UNSUPPORTED, "padded dims above 512 (f32) / 1024 (bf16) / 2048 "
                                           "(fp8) are not supported by the MFMA scan");
    }
}

uint32_t cw_for(uint32_t k) {
    uint32_t cw = 128;
    while (cw < k + 64) cw <<= 1;
    return cw;
}

void launch_sample_threshold(const float *scores, uint32_t nq, uint32_t n, uint32_t ld,
                             uint32_t kp, const float *delta, unsigned long long *thresh,
                             float *thresh_f, hipStream_t s) {
    switch (cw_for(kp)) {
    case 128:
        hipLaunchKernelGGL(sample_threshold_kernel<128>, dim3(nq), dim3(256), 0, s, scores, n, ld,
                           kp, delta, thresh, thresh_f);
        break;
    case 256:
        hipLaunchKernelGGL(sample_threshold_kernel<256>, dim3(nq), dim3(256), 0, s, scores, n, ld,
                           kp, delta, thresh, thresh_f);
        break;
    case 512:
        hipLaunchKernelGGL(sample_threshold_kernel<512>, dim3(nq), dim3(256), 0, s, scores, n, ld,
                           kp, delta, thresh, thresh_f);
        break;
    default:
        throw Error(NNDEX_ERR_UNSUPPORTED, "over-fetch above 448 is not supported");
    }
    NNDEX_HIP_CHECK(hipGetLastError());
}

void launch_rerank_full(const RerankArgs &a, uint32_t nq, hipStream_t s) {
    switch (cw_for(a.k)) {
    case 128:
        hipLaunchKernelGGL(rerank_full_kernel<128>, dim3(nq), dim3(256), 0, s, a);
        break;
    case 256:
        hipLaunchKernelGGL(rerank_full_kernel<256>, dim3(nq), dim3(256), 0, s, a);
        break;
    case 512:
        hipLaunchKernelGGL(rerank_full_kernel<512>, dim3(nq), dim3(256), 0, s, a);
        break;
    default:
        return;  // k above 448: queries stay flagged for the exact scan
    }
    NNDEX_HIP_CHECK(hipGetLastError());
}

void launch_rerank(const RerankArgs &a, uint32_t nq, hipStream_t s) {
    switch (cw_for(a.kp + 1)) {
    case 128:
        hipLaunchKernelGGL(rerank_kernel<128>, dim3(nq), dim3(256), 0, s, a);
        break;
    case 256:
        hipLaunchKernelGGL(rerank_kernel<256>, dim3(nq), dim3(256), 0, s, a);
        break;
    case 512:
        hipLaunchKernelGGL(rerank_kernel<512>, dim3(nq), dim3(256), 0, s, a);
        break;
    default:
        throw Error(NNDEX_ERR_UNSUPPORTED, "over-fetch above 447 is not supported");
    }
    NNDEX_HIP_CHECK(hipGetLastError());
}

}  // namespace nndex
