// Batched search on MFMA with an exact FP32 rerank and a per-query
// certificate.
//
// Replaces the batch path of the reference GPU backend (src/gpu.rs:595-785:
// one WGSL dispatch per (row chunk, query sub-batch), full Q x rows score
// readback, a single-threaded host heap per query) with four device stages:
//
//  1. sample threshold: the batch is scored (same MFMA kernel) against a
//     strided sample of the rows; the k'-th best sample score, lowered by the
//     per-query error bound delta, is a threshold every true top-k' row passes;
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
#include "common.hpp"
#include "wave_topk.hpp"

namespace nndex {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// Element traits: one 16-byte unit per lane per MFMA step group.
// ---------------------------------------------------------------------------
struct F32E {
    // four 16x16x4 f32 MFMAs over the unit's four floats (k = 4 each)
    static __device__ inline v4f mma(uint4 a, uint4 b, v4f c) {
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(__builtin_bit_cast(float, a.x),
                                                 __builtin_bit_cast(float, b.x), c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(__builtin_bit_cast(float, a.y),
                                                 __builtin_bit_cast(float, b.y), c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(__builtin_bit_cast(float, a.z),
                                                 __builtin_bit_cast(float, b.z), c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(__builtin_bit_cast(float, a.w),
                                                 __builtin_bit_cast(float, b.w), c, 0, 0, 0);
        return c;
    }
};
struct BF16E {
    // one 16x16x32 bf16 MFMA over the unit's eight bf16 values
    static __device__ inline v4f mma(uint4 a, uint4 b, v4f c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a),
                                                       __builtin_bit_cast(v8bf, b), c, 0, 0, 0);
    }
};

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

// Largest row norm of the tiled matrix, in f64 (bounds |x| in delta).
__global__ __launch_bounds__(256) void max_norm_kernel(const float4 *__restrict__ x,
                                                       uint32_t rows, uint32_t d4,
                                                       unsigned long long *__restrict__ out) {
    uint32_t row = blockIdx.x * 256 + threadIdx.x;
    double ss = 0.0;
    if (row < rows) {
        const float4 *b = x + (size_t)(row / 64) * d4 * 64 + row % 64;
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
// Query preparation: low-precision copy (row-major, 16-byte units) and the
// per-query error bound delta of |approximate - exact| over any row:
//   exact side   gamma_n |x| |q|   (four-lane FMA chain, n = padded dims)
//   approx side  gamma_n |x~| |q~| (MFMA accumulation)
//   rounding     |x~| |q~ - q| + u_e |x| |q|   (u_e = 2^-8 for bf16, 0 for f32)
// with |x| <= xmax (largest row norm) and |x~| <= (1 + u_e) xmax.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void prep_queries_kernel(const float *__restrict__ q,
                                                           uint32_t nq, uint32_t pdims,
                                                           int bf16, uint4 *__restrict__ qlow,
                                                           const unsigned long long *xmax_sq,
                                                           float *__restrict__ delta) {
    const uint32_t qi = blockIdx.x;
    if (qi >= nq) return;
    const float *src = q + (size_t)qi * pdims;
    double nq2 = 0, nl2 = 0, nd2 = 0;
    for (uint32_t i = threadIdx.x; i < pdims; i += 256) {
        double v = src[i];
        double lv = bf16 ? (double)(float)(__bf16)src[i] : v;
        nq2 += v * v;
        nl2 += lv * lv;
        nd2 += (lv - v) * (lv - v);
    }
    if (bf16) {
        uint32_t d8 = pdims / 8;
        for (uint32_t c = threadIdx.x; c < d8; c += 256) {
            const float *s = src + c * 8;
            qlow[(size_t)qi * d8 + c] = make_uint4(pack_bf16x2(s[0], s[1]), pack_bf16x2(s[2], s[3]),
                                                   pack_bf16x2(s[4], s[5]), pack_bf16x2(s[6], s[7]));
        }
    }
    __shared__ double red[3][256];
    red[0][threadIdx.x] = nq2;
    red[1][threadIdx.x] = nl2;
    red[2][threadIdx.x] = nd2;
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int j = 0; j < 3; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double xm = sqrt(__builtin_bit_cast(double, *xmax_sq));
        double ue = bf16 ? 0x1p-8 : 0.0;
        double u = 0x1p-24, n = pdims;
        double gamma = n * u / (1.0 - n * u);
        double qn = sqrt(red[0][0]), ql = sqrt(red[1][0]), qd = sqrt(red[2][0]);
        double xl = xm * (1.0 + ue);
        double d = gamma * xm * qn + gamma * xl * ql + xl * qd + ue * xm * qn;
        delta[qi] = (float)(d * (1.0 + 1e-6)) + 1e-30f;
    }
}

// ---------------------------------------------------------------------------
// X-stationary MFMA scan.  NW waves per block, RW rows per wave (multiple of
// 16), MAXS 16-byte units per lane per row (units per row = 4 * steps; lane
// group g = lane/16 owns units 4s+g).  Mode 0 appends candidates, mode 1
// writes scores.
// ---------------------------------------------------------------------------
template <typename E, int NW, int RW, int MAXS>
__global__ __launch_bounds__(NW * 64) void mfma_scan_kernel(MfmaScanArgs a) {
    constexpr int MT = RW / 16;  // 16-row MFMA tiles per wave
    extern __shared__ uint4 qlds[];  // 2 buffers x [4*MAXS units][17 slots]
    constexpr uint32_t kStride = 17;
    const uint32_t lane = __lane_id();
    const uint32_t wid = threadIdx.x >> 6;
    const uint32_t g = lane >> 4, l16 = lane & 15;
    const uint32_t dch = a.dch;
    const uint32_t steps = (dch + 3) / 4;
    const uint32_t units_pad = steps * 4;
    const uint32_t buf_units = units_pad * kStride;
    const uint32_t block_rows = NW * RW;
    const uint32_t nblk = (a.rows_padded + block_rows - 1) / block_rows;
    const uint32_t nchunks = (a.nq + 15) / 16;

    // zero the padding units of both buffers once (chunks >= dch)
    for (uint32_t i = threadIdx.x; i < 2 * buf_units; i += NW * 64) {
        uint32_t unit = (i % buf_units) / kStride;
        if (unit >= dch) qlds[i] = make_uint4(0, 0, 0, 0);
    }

    for (uint32_t bt = blockIdx.x; bt < nblk; bt += gridDim.x) {
        // ---- stationary X: MT x MAXS units per lane ----
        uint4 xr[MT][MAXS];
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const uint32_t row = bt * block_rows + wid * RW + m * 16 + l16;
            const bool in = row < a.rows_padded;
            const uint4 *base = a.x + (size_t)(row / 64) * dch * 64 + (row % 64);
#pragma unroll
            for (int s = 0; s < MAXS; ++s) {
                const uint32_t c = 4 * s + g;
                xr[m][s] = (in && s < (int)steps && c < dch) ? base[(size_t)c * 64]
                                                             : make_uint4(0, 0, 0, 0);
            }
        }
        // ---- stream query chunks through LDS (double-buffered) ----
        auto load_chunk = [&](uint32_t ch, uint32_t buf) {
            uint4 *dst = qlds + buf * buf_units;
            for (uint32_t i = threadIdx.x; i < 16 * dch; i += NW * 64) {
                uint32_t n = i / dch, c = i % dch;
                uint32_t qi = ch * 16 + n;
                dst[c * kStride + n] =
                    qi < a.nq ? a.q[(size_t)qi * a.q_stride + c] : make_uint4(0, 0, 0, 0);
            }
        };
        __syncthreads();
        load_chunk(0, 0);
        __syncthreads();
        for (uint32_t ch = 0; ch < nchunks; ++ch) {
            const uint32_t buf = ch & 1;
            if (ch + 1 < nchunks) load_chunk(ch + 1, buf ^ 1);
            const uint4 *src = qlds + buf * buf_units;
            v4f acc[MT];
#pragma unroll
            for (int m = 0; m < MT; ++m) acc[m] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < MAXS; ++s) {
                if (s < (int)steps) {
                    uint4 qv = src[(4 * s + g) * kStride + l16];
#pragma unroll
                    for (int m = 0; m < MT; ++m) acc[m] = E::mma(xr[m][s], qv, acc[m]);
                }
            }
            // ---- epilogue: lane holds rows 4*g+v of each tile, query l16 ----
            const uint32_t qi = ch * 16 + l16;
            if (a.mode == 0) {
                unsigned long long th = qi < a.nq ? a.thresh[qi] : ~0ull;
#pragma unroll
                for (int m = 0; m < MT; ++m) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const uint32_t row = bt * block_rows + wid * RW + m * 16 + 4 * g + v;
                        const uint64_t key = make_key(acc[m][v], row);
                        const bool pass = qi < a.nq && row < a.rows && key >= th;
                        if (__ballot(pass) != 0 && pass) {
                            uint32_t pos = atomicAdd(a.cand_count + qi, 1u);
                            if (pos < a.capq) a.cand[(size_t)qi * a.capq + pos] = key;
                        }
                    }
                }
            } else {
#pragma unroll
                for (int m = 0; m < MT; ++m)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const uint32_t row = bt * block_rows + wid * RW + m * 16 + 4 * g + v;
                        if (qi < a.nq && row < a.rows) a.out[(size_t)qi * a.ldo + row] = acc[m][v];
                    }
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// Sample threshold: per query, the k'-th best of `n` sample scores, lowered
// by delta, as a key threshold (keys >= thresh pass).  n < k' -> 0 (all pass).
// ---------------------------------------------------------------------------
template <int CW>
__global__ __launch_bounds__(256) void sample_threshold_kernel(const float *__restrict__ scores,
                                                               uint32_t n, uint32_t ld,
                                                               uint32_t kp,
                                                               const float *__restrict__ delta,
                                                               unsigned long long *thresh) {
    __shared__ uint64_t slots[4][CW];
    __shared__ uint32_t cnt[4];
    const uint32_t q = blockIdx.x, lane = __lane_id(), wid = threadIdx.x >> 6;
    if (n < kp) {
        if (threadIdx.x == 0) thresh[q] = 0;
        return;
    }
    WaveTopK tk;
    tk.init(slots[wid], CW, kp, nullptr);
    const float *s = scores + (size_t)q * ld;
    for (uint32_t i0 = wid * 64; i0 < n; i0 += 256) {
        uint32_t i = i0 + lane;
        tk.offer(i < n ? make_key(s[i], i) : 0, i < n);
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
            float t = key_sim(slots[0][kp - 1]);
            float lowered = (float)((double)t - (double)delta[q]);
            // round toward -inf so the float threshold never exceeds t - delta
            if ((double)lowered > (double)t - (double)delta[q]) {
                // one ulp toward -inf
                uint32_t b = __builtin_bit_cast(uint32_t, lowered);
                if (lowered > 0.f) b -= 1;
                else if (lowered < 0.f) b += 1;
                else b = 0x80000001u;
                lowered = __builtin_bit_cast(float, b);
            }
            thresh[q] = (unsigned long long)ordered_bits(lowered) << 32;
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
    const uint32_t q = blockIdx.x, lane = __lane_id(), wid = threadIdx.x >> 6;
    const uint32_t total = a.cand_count[q];
    const bool overflow = total > a.capq;
    const uint32_t n = overflow ? a.capq : total;
    const uint32_t kp1 = a.kp + 1;
    const uint64_t *c = a.cand + (size_t)q * a.capq;

    WaveTopK tk;
    tk.init(slots[wid], CW, kp1, nullptr);
    for (uint32_t i0 = wid * 64; i0 < n; i0 += 256) {
        uint32_t i = i0 + lane;
        tk.offer(i < n ? c[i] : 0, i < n);
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
    // exact rescoring, one thread per candidate, in the reference dot order
    for (uint32_t i = threadIdx.x; i < CW; i += 256) {
        uint64_t key = 0;
        if (i < nres) {
            const uint32_t row = key_row(slots[0][i]);
            const float4 *xb = a.x + (size_t)(row / 64) * a.d4 * 64 + row % 64;
            const float4 *qb = a.q + (size_t)q * a.d4;
            float l0 = 0.f, l1 = 0.f, l2 = 0.f, l3 = 0.f;
            for (uint32_t cc = 0; cc < a.d4; ++cc) {
                float4 xv = xb[(size_t)cc * 64], qv = qb[cc];
                l0 = __builtin_fmaf(xv.x, qv.x, l0);
                l1 = __builtin_fmaf(xv.y, qv.y, l1);
                l2 = __builtin_fmaf(xv.z, qv.z, l2);
                l3 = __builtin_fmaf(xv.w, qv.w, l3);
            }
            key = make_key((l0 + l1) + (l2 + l3), row);
        }
        exact[i] = key;
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

void launch_to_bf16(const float4 *x, size_t rows_padded, uint32_t d8, uint4 *y, hipStream_t s) {
    size_t units = rows_padded * d8;
    uint32_t grid = (uint32_t)((units + 255) / 256);
    if (grid > 65535u * 8) grid = 65535u * 8;
    hipLaunchKernelGGL(to_bf16_kernel, dim3(grid), dim3(256), 0, s, x, units, d8, y);
    NNDEX_HIP_CHECK(hipGetLastError());
}

void launch_max_norm(const float4 *x, uint32_t rows, uint32_t d4, unsigned long long *out,
                     hipStream_t s) {
    hipLaunchKernelGGL(max_norm_kernel, dim3((rows + 255) / 256), dim3(256), 0, s, x, rows, d4,
                       out);
    NNDEX_HIP_CHECK(hipGetLastError());
}

void launch_prep_queries(const float *q, uint32_t nq, uint32_t pdims, bool bf16, uint4 *qlow,
                         const unsigned long long *xmax_sq, float *delta, hipStream_t s) {
    hipLaunchKernelGGL(prep_queries_kernel, dim3(nq), dim3(256), 0, s, q, nq, pdims, bf16 ? 1 : 0,
                       qlow, xmax_sq, delta);
    NNDEX_HIP_CHECK(hipGetLastError());
}

template <typename E, int NW, int RW, int MAXS>
static void launch_mfma_t(const MfmaScanArgs &a, uint32_t grid, hipStream_t s) {
    const uint32_t steps = (a.dch + 3) / 4;
    const size_t lds = 2ull * steps * 4 * 17 * sizeof(uint4);
    auto fn = mfma_scan_kernel<E, NW, RW, MAXS>;
    NNDEX_HIP_CHECK(hipFuncSetAttribute((const void *)fn,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(fn, dim3(grid), dim3(NW * 64), lds, s, a);
    NNDEX_HIP_CHECK(hipGetLastError());
}

uint32_t mfma_block_rows(bool bf16) { return bf16 ? 8 * 32 : 4 * 16; }

void launch_mfma_scan(const MfmaScanArgs &a, bool bf16, uint32_t grid, hipStream_t s) {
    const uint32_t steps = (a.dch + 3) / 4;
    if (steps > 32)
        throw Error(NNDEX_ERR_UNSUPPORTED, "padded dims above 512 (f32) / 1024 (bf16) are not "
                                           "supported by the MFMA scan");
    if (!bf16) {
        if (steps <= 8) return launch_mfma_t<F32E, 4, 16, 8>(a, grid, s);
        if (steps <= 16) return launch_mfma_t<F32E, 4, 16, 16>(a, grid, s);
        if (steps <= 24) return launch_mfma_t<F32E, 4, 16, 24>(a, grid, s);
        return launch_mfma_t<F32E, 4, 16, 32>(a, grid, s);
    }
    if (steps <= 8) return launch_mfma_t<BF16E, 8, 32, 8>(a, grid, s);
    if (steps <= 16) return launch_mfma_t<BF16E, 8, 32, 16>(a, grid, s);
    if (steps <= 24) return launch_mfma_t<BF16E, 8, 16, 24>(a, grid, s);
    return launch_mfma_t<BF16E, 8, 16, 32>(a, grid, s);
}

uint32_t cw_for(uint32_t k) {
    uint32_t cw = 128;
    while (cw < k + 64) cw <<= 1;
    return cw;
}

void launch_sample_threshold(const float *scores, uint32_t nq, uint32_t n, uint32_t ld,
                             uint32_t kp, const float *delta, unsigned long long *thresh,
                             hipStream_t s) {
    switch (cw_for(kp)) {
    case 128:
        hipLaunchKernelGGL(sample_threshold_kernel<128>, dim3(nq), dim3(256), 0, s, scores, n, ld,
                           kp, delta, thresh);
        break;
    case 256:
        hipLaunchKernelGGL(sample_threshold_kernel<256>, dim3(nq), dim3(256), 0, s, scores, n, ld,
                           kp, delta, thresh);
        break;
    case 512:
        hipLaunchKernelGGL(sample_threshold_kernel<512>, dim3(nq), dim3(256), 0, s, scores, n, ld,
                           kp, delta, thresh);
        break;
    default:
        throw Error(NNDEX_ERR_UNSUPPORTED, "over-fetch above 448 is not supported");
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
