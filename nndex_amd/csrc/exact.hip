// Exact FP32 search kernels for gfx950: layout transform, row L2
// normalisation, and the fused score + top-k scan.
//
// Reference being replaced: the WGSL dot kernels (src/shaders/dot_lowdim.wgsl,
// dot_highdim.wgsl), the full-score readback and host heap top-k of
// src/gpu.rs:511-592 / :640-785, and the host normalisation of
// src/lib.rs:461-575.
//
// Arithmetic follows the reference CPU path exactly: every score is the
// four-lane f32 FMA chain of simsimd's dot kernel (lane j accumulates elements
// 4c+j for c = 0, 1, 2, ... in order, then (l0+l1)+(l2+l3)), so scores are
// bit-identical to the CPU restatement of src/cpu.rs:653-660.
#include "common.hpp"
#include "kernels.hpp"
#include "wave_topk.hpp"

namespace nndex {

// ---------------------------------------------------------------------------
// Layout transform: row-major [rows][dims] -> tiled float4 layout with zero
// padding of columns to d4*4 and of rows to a multiple of 64.  Sets bit 0 of
// *flags if any input value is not finite (ensure_finite, src/lib.rs:445-450).
// One block handles one 64-row tile; the tile is staged in LDS in column
// chunks of 64 floats so global reads and writes are both contiguous.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void to_tiled_kernel(const float *__restrict__ src,
                                                       uint32_t rows, uint32_t dims, uint32_t d4,
                                                       float4 *__restrict__ dst,
                                                       uint32_t *__restrict__ flags) {
    __shared__ float stage[64][64 + 1];
    const uint32_t tile = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    bool bad = false;
    for (uint32_t c0 = 0; c0 < d4 * 4; c0 += 64) {
        // load: 64 rows x 64 columns, consecutive threads take consecutive columns
        for (uint32_t i = tid; i < 64 * 64; i += 256) {
            uint32_t r = i / 64, c = i % 64;
            size_t row = (size_t)tile * 64 + r;
            uint32_t col = c0 + c;
            float v = 0.0f;
            if (row < rows && col < dims) {
                v = src[row * dims + col];
                bad |= !__builtin_isfinite(v);
            }
            stage[r][c] = v;
        }
        __syncthreads();
        // store: for each float4 chunk, 64 rows contiguous (1 KiB)
        for (uint32_t i = tid; i < 16 * 64; i += 256) {
            uint32_t c4l = i / 64, r = i % 64;
            uint32_t c4 = c0 / 4 + c4l;
            if (c4 < d4) {
                float4 v = make_float4(stage[r][c4l * 4 + 0], stage[r][c4l * 4 + 1],
                                       stage[r][c4l * 4 + 2], stage[r][c4l * 4 + 3]);
                dst[((size_t)tile * d4 + c4) * 64 + r] = v;
            }
        }
        __syncthreads();
    }
    if (__ballot(bad) != 0 && __lane_id() == 0) atomicOr(flags, 1u);
}

// Squared norm with the reference's rules (src/lib.rs:519-553):
//   dims >= 16: simsimd dot of the row with itself over its `dims` values:
//               four FMA lanes over full chunks, (l0+l1)+(l2+l3), then the
//               dims%4 tail elements FMA'd sequentially;
//   dims <  16: sequential f32 `sq += v*v` without fusion.
// `get(c4)` returns float4 chunk c4 of the row.
template <typename Get>
__device__ inline float norm_sq_ref(Get get, uint32_t simd_dims, uint32_t dims, bool &nonfinite) {
    if (simd_dims >= 16) {
        float l0 = 0.f, l1 = 0.f, l2 = 0.f, l3 = 0.f;
        uint32_t full = dims / 4;
        for (uint32_t c = 0; c < full; ++c) {
            float4 v = get(c);
            l0 = __builtin_fmaf(v.x, v.x, l0);
            l1 = __builtin_fmaf(v.y, v.y, l1);
            l2 = __builtin_fmaf(v.z, v.z, l2);
            l3 = __builtin_fmaf(v.w, v.w, l3);
        }
        float ab = (l0 + l1) + (l2 + l3);
        uint32_t tail = dims % 4;
        if (tail) {
            float4 v = get(full);
            ab = __builtin_fmaf(v.x, v.x, ab);
            if (tail > 1) ab = __builtin_fmaf(v.y, v.y, ab);
            if (tail > 2) ab = __builtin_fmaf(v.z, v.z, ab);
        }
        nonfinite = !__builtin_isfinite(ab);
        return ab;
    }
    float sq = 0.f;
    bool bad = false;
    for (uint32_t c = 0; c * 4 < dims; ++c) {
        float4 v = get(c);
        float e[4] = {v.x, v.y, v.z, v.w};
        for (uint32_t j = 0; j < 4 && c * 4 + j < dims; ++j) {
            bad |= !__builtin_isfinite(e[j]);
            sq = __fadd_rn(sq, __fmul_rn(e[j], e[j]));
        }
    }
    nonfinite = bad;
    return sq;
}

// Row normalisation in the tiled layout, one thread per row, in place.
// Flags: bit 0 non-finite (NNdexError::NonFiniteInput), bit 1 zero norm
// (NNdexError::ZeroNorm: norm_sq <= f32::EPSILON).
__global__ __launch_bounds__(256) void normalize_tiled_kernel(float4 *__restrict__ x,
                                                              uint32_t rows, uint32_t dims,
                                                              uint32_t d4,
                                                              uint32_t *__restrict__ flags) {
    const uint32_t row = blockIdx.x * 256 + threadIdx.x;
    uint32_t f = 0;
    if (row < rows) {
        float4 *base = x + (size_t)(row / 64) * d4 * 64 + (row % 64);
        bool nonfinite = false;
        float sq = norm_sq_ref([&](uint32_t c) { return base[(size_t)c * 64]; }, dims, dims,
                               nonfinite);
        if (nonfinite) {
            f = 1;
        } else if (sq <= 1.1920929e-7f) {
            f = 2;
        } else {
            float inv = 1.0f / __builtin_sqrtf(sq);  // both correctly rounded (IEEE), as Rust's sqrt().recip()
            for (uint32_t c = 0; c < d4; ++c) {
                float4 v = base[(size_t)c * 64];
                v.x = __fmul_rn(v.x, inv);
                v.y = __fmul_rn(v.y, inv);
                v.z = __fmul_rn(v.z, inv);
                v.w = __fmul_rn(v.w, inv);
                base[(size_t)c * 64] = v;
            }
        }
    }
    uint32_t any1 = __ballot(f & 1) != 0, any2 = __ballot(f & 2) != 0;
    if (__lane_id() == 0 && (any1 || any2)) atomicOr(flags, (any1 ? 1u : 0u) | (any2 ? 2u : 0u));
}

// Row-major query preparation: src [rows][dims] -> dst [rows][pdims].
//   mode 0 (kQueryPad):    ensure_finite + pad_rows / pad_vector
//                          (src/lib.rs:383-385, :405-408, :578-599)
//   mode 1 (kQueryBatch):  normalize_rows (src/lib.rs:461-553): the simsimd
//                          four-lane sum for dims >= 16, sequential below
//   mode 2 (kQuerySingle): normalize_vector (src/lib.rs:558-575): always the
//                          sequential unfused f32 sum of squares
// A row that fails (bit 0 non-finite, bit 1 zero norm) is written as zeros;
// the bits go to *flags (OR over rows) and, when given, row_flags[row].
__global__ __launch_bounds__(256) void normalize_rowmajor_kernel(
    const float *__restrict__ src, uint32_t rows, uint32_t dims, uint32_t pdims, int mode,
    float *__restrict__ dst, uint32_t *__restrict__ flags, uint8_t *__restrict__ row_flags) {
    const uint32_t row = blockIdx.x * 256 + threadIdx.x;
    uint32_t f = 0;
    if (row < rows) {
        const float *s = src + (size_t)row * dims;
        float *d = dst + (size_t)row * pdims;
        auto get = [&](uint32_t c) {
            float e[4];
            for (uint32_t j = 0; j < 4; ++j) e[j] = c * 4 + j < dims ? s[c * 4 + j] : 0.f;
            return make_float4(e[0], e[1], e[2], e[3]);
        };
        if (mode != kQueryPad) {
            bool nonfinite = false;
            // the single-vector form never takes the four-lane sum
            const float sq = norm_sq_ref(get, mode == kQuerySingle ? 0u : dims, dims, nonfinite);
            if (nonfinite) {
                f = 1;
            } else if (sq <= 1.1920929e-7f) {
                f = 2;
            } else {
                float inv = 1.0f / __builtin_sqrtf(sq);  // both correctly rounded (IEEE), as Rust's sqrt().recip()
                for (uint32_t i = 0; i < dims; ++i) d[i] = __fmul_rn(s[i], inv);
            }
        } else {
            for (uint32_t i = 0; i < dims; ++i) {
                float v = s[i];
                if (!__builtin_isfinite(v)) f = 1;
                d[i] = v;
            }
        }
        if (f)
            for (uint32_t i = 0; i < dims; ++i) d[i] = 0.f;
        for (uint32_t i = dims; i < pdims; ++i) d[i] = 0.f;
        if (row_flags) row_flags[row] = static_cast<uint8_t>(f);
    }
    uint32_t any1 = __ballot(f & 1) != 0, any2 = __ballot(f & 2) != 0;
    if (__lane_id() == 0 && (any1 || any2)) atomicOr(flags, (any1 ? 1u : 0u) | (any2 ? 2u : 0u));
}

// ---------------------------------------------------------------------------
// Fused exact scan: scores of QB queries against every row + top-k.
//
// Each wave owns a sequence of 64-row tiles; lane r computes the scores of
// row r of the tile by walking its float4 chunks in order (coalesced 1 KiB
// loads per chunk; query chunks are wave-uniform and come through the scalar
// cache).  Scores become ranking keys offered to a per-wave WaveTopK.  At the
// end the block's four waves are merged, the block writes its sorted top-k to
// `part`, and the last block to finish (ticket) merges all block lists,
// writes the outputs and resets the ticket and thresholds for the next launch.
// ---------------------------------------------------------------------------
template <int QB, int CW, int U, bool NT>
__global__ __launch_bounds__(256) void exact_scan_kernel(ScanArgs a) {
    __shared__ uint64_t slots[4][QB][CW];
    __shared__ uint32_t wave_count[4][QB];
    __shared__ uint32_t is_last;
    __shared__ uint32_t pool_n, sel_hi;

    const uint32_t lane = __lane_id();
    const uint32_t wid = threadIdx.x >> 6;
    const uint32_t d4 = a.d4;
    const uint32_t ntiles = (a.rows + 63) / 64;
    const float4 *__restrict__ X = a.x;
    const_as<f32x4> *Q = as_constant(reinterpret_cast<const f32x4 *>(a.q));

    // Per-wave thresholds only: a device-wide one costs more than it prunes
    // here, since reading it back waits for every outstanding row load of
    // the wave (vector memory counters complete in order).  Blocks publish
    // their merged lists to the final merge.
    WaveTopK tk[QB];
#pragma unroll
    for (int q = 0; q < QB; ++q) tk[q].init(slots[wid][q], CW, a.k, nullptr, a.k_out);

    // U float4 chunks per load group (two groups in flight); NT selects
    // streaming (non-temporal) loads at compile time so the loop is straight
    auto load_group = [&](float4(&buf)[U], const float4 *p, uint32_t cbase) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if constexpr (NT)
                buf[u] = load_nt(p + (size_t)(cbase + u) * 64);
            else
                buf[u] = p[(size_t)(cbase + u) * 64];
        }
    };
    auto fma_group = [&](float(&acc)[QB][4], const float4(&buf)[U], uint32_t cbase) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int q = 0; q < QB; ++q) {
                const uint32_t qq = (uint32_t)q < a.nvalid ? (uint32_t)q : a.nvalid - 1;
                const f32x4 qv = Q[(size_t)qq * a.q_stride4 + cbase + u];
                acc[q][0] = __builtin_fmaf(buf[u].x, qv.x, acc[q][0]);
                acc[q][1] = __builtin_fmaf(buf[u].y, qv.y, acc[q][1]);
                acc[q][2] = __builtin_fmaf(buf[u].z, qv.z, acc[q][2]);
                acc[q][3] = __builtin_fmaf(buf[u].w, qv.w, acc[q][3]);
            }
        }
    };
    auto finish_tile = [&](float(&acc)[QB][4], uint32_t t) {
        const uint32_t row = t * 64 + lane;
        const bool valid = row < a.rows;
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            float s = (acc[q][0] + acc[q][1]) + (acc[q][2] + acc[q][3]);
            tk[q].offer(make_key(s, row), valid);
        }
    };

    // Tile order: interleaved (wave w takes w, w + W, ...) or contiguous
    // (wave w takes a run of ceil(ntiles / W) consecutive tiles).
    const uint32_t nwaves = gridDim.x * 4;
    const uint32_t wg = blockIdx.x * 4 + wid;
    uint32_t t, tend, tstride;
    if (a.flags & kScanContiguous) {
        const uint32_t run = (ntiles + nwaves - 1) / nwaves;
        t = wg * run;
        tend = t + run < ntiles ? t + run : ntiles;
        tstride = 1;
    } else {
        t = wg;
        tend = ntiles;
        tstride = nwaves;
    }
    if (d4 % (2 * U) == 0) {
        // Two load groups in flight: group g+1 (or the next tile's first
        // group) is issued before group g is consumed, so the wave keeps
        // U-2U KiB of reads outstanding across tile boundaries.
        float4 A[U], B[U];
        const float4 *p = X + (size_t)t * d4 * 64 + lane;
        if (t < tend) load_group(A, p, 0);  // (wave-uniform)
        while (t < tend) {
            const uint32_t tn = t + tstride;
            const float4 *pn = X + (size_t)tn * d4 * 64 + lane;
            float acc[QB][4];
#pragma unroll
            for (int q = 0; q < QB; ++q) acc[q][0] = acc[q][1] = acc[q][2] = acc[q][3] = 0.f;
            // the chunk loop without its last pair keeps a straight body; the
            // last pair prefetches the next tile's first group (if any)
            uint32_t c = 0;
            for (; c + 2 * U < d4; c += 2 * U) {
                load_group(B, p, c + U);
                fma_group(acc, A, c);
                load_group(A, p, c + 2 * U);
                fma_group(acc, B, c + U);
            }
            load_group(B, p, c + U);
            fma_group(acc, A, c);
            if (tn < tend) load_group(A, pn, 0);  // (wave-uniform)
            fma_group(acc, B, c + U);
            finish_tile(acc, t);
            t = tn;
            p = pn;
        }
    } else {
        for (; t < tend; t += tstride) {
            float acc[QB][4];
#pragma unroll
            for (int q = 0; q < QB; ++q) acc[q][0] = acc[q][1] = acc[q][2] = acc[q][3] = 0.f;
            const float4 *xt = X + (size_t)t * d4 * 64 + lane;
#pragma unroll 8
            for (uint32_t c = 0; c < d4; ++c) {
                float4 v;
                if constexpr (NT)
                    v = load_nt(xt + (size_t)c * 64);
                else
                    v = xt[(size_t)c * 64];
#pragma unroll
                for (int q = 0; q < QB; ++q) {
                    const uint32_t qq = (uint32_t)q < a.nvalid ? (uint32_t)q : a.nvalid - 1;
                    const f32x4 qv = Q[(size_t)qq * a.q_stride4 + c];
                    acc[q][0] = __builtin_fmaf(v.x, qv.x, acc[q][0]);
                    acc[q][1] = __builtin_fmaf(v.y, qv.y, acc[q][1]);
                    acc[q][2] = __builtin_fmaf(v.z, qv.z, acc[q][2]);
                    acc[q][3] = __builtin_fmaf(v.w, qv.w, acc[q][3]);
                }
            }
            finish_tile(acc, t);
        }
    }

    // Block merge: waves 1..3 hand their lists to wave 0.
#pragma unroll
    for (int q = 0; q < QB; ++q) {
        tk[q].compact();
        if (lane == 0) wave_count[wid][q] = tk[q].count;
    }
    __syncthreads();
    if (wid == 0) {
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            for (uint32_t w = 1; w < 4; ++w) {
                uint32_t n = wave_count[w][q];
                for (uint32_t i0 = 0; i0 < n; i0 += 64) {
                    uint32_t i = i0 + lane;
                    uint64_t key = i < n ? slots[w][q][i] : 0;
                    tk[q].offer(key, i < n);
                }
            }
            tk[q].compact();
            // Publish the block's list (sorted, at most k keys) in the
            // block's own k slots of the query's table, empty slots as key 0
            // (no key is 0), so the final merge finds each block's largest
            // key at the start of its slots
            uint64_t *dst = a.part + ((size_t)q * gridDim.x + blockIdx.x) * a.k;
            for (uint32_t i = lane; i < a.k; i += 64) dst[i] = i < tk[q].count ? slots[0][q][i] : 0;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        uint32_t ticket = 0;
        if (lane == 0)
            ticket = __hip_atomic_fetch_add(a.ticket, 1u, __ATOMIC_ACQ_REL,
                                            __HIP_MEMORY_SCOPE_AGENT);
        ticket = __shfl(ticket, 0);
        if (lane == 0) is_last = ticket == gridDim.x - 1;
    }
    __syncthreads();
    if (!is_last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");

    // Final merge, one query at a time.  (1) S = the k_out-th largest score
    // among the blocks' largest keys, found bit by bit with ballots: at least
    // k_out rows score S or more, so a row scoring below S is neither in the
    // result nor tied with its last score.  (2) Every published key scoring
    // at least S goes to one LDS pool (the slots, free now).  (3) Wave 0
    // selects the result from the pool as a wave's list would.  A pool too
    // small for the keys that pass (ties or large k) falls back to every
    // wave offering its share of the table; more than 512 blocks skip (1).
    const uint32_t nb = gridDim.x;
    uint64_t *pool = &slots[0][0][0];
    constexpr uint32_t POOL = 4 * QB * CW;
    constexpr int MB = 8;   // block maxima per lane in (1)
    constexpr int KB = 16;  // table keys per lane per batch
#pragma unroll
    for (int q = 0; q < QB; ++q) {
        const uint64_t *tab = a.part + (size_t)q * nb * a.k;
        const uint32_t nslots = nb * a.k;
        // the first batch of table keys is in flight during (1)
        uint64_t key[KB];
#pragma unroll
        for (int u = 0; u < KB; ++u) {
            const uint32_t j = (u * 4 + wid) * 64 + lane;
            key[u] = j < nslots ? tab[j] : 0;
        }
        if (wid == 0) {
            uint32_t sel = 0;  // wave-uniform
            if (nb >= a.k_out && nb <= 64 * MB) {
                uint32_t hi[MB];
#pragma unroll
                for (int i = 0; i < MB; ++i) {
                    const uint32_t b = i * 64 + lane;
                    hi[i] = b < nb ? static_cast<uint32_t>(tab[(size_t)b * a.k] >> 32) : 0;
                }
                for (uint32_t bit = 32; bit-- > 0;) {
                    const uint32_t cand = sel | (1u << bit);
                    uint32_t cnt = 0;
#pragma unroll
                    for (int i = 0; i < MB; ++i) cnt += __popcll(__ballot(hi[i] >= cand));
                    if (cnt >= a.k_out) sel = cand;
                }
            }
            if (lane == 0) {
                sel_hi = sel;
                pool_n = 0;
            }
        }
        __syncthreads();
        const uint64_t floor_key = static_cast<uint64_t>(sel_hi) << 32;
        for (uint32_t j0 = 0;; j0 += 256 * KB) {
#pragma unroll
            for (int u = 0; u < KB; ++u) {
                const uint32_t j = j0 + (u * 4 + wid) * 64 + lane;
                const bool pass = j < nslots && key[u] != 0 && key[u] >= floor_key;
                const uint64_t m = __ballot(pass);
                if (m == 0) continue;
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(&pool_n, static_cast<uint32_t>(__popcll(m)));
                base = __shfl(base, 0);
                const uint32_t pos = base + __popcll(m & __lanemask_lt());
                if (pass && pos < POOL) pool[pos] = key[u];
            }
            if (j0 + 256 * KB >= nslots) break;
#pragma unroll
            for (int u = 0; u < KB; ++u) {
                const uint32_t j = j0 + 256 * KB + (u * 4 + wid) * 64 + lane;
                key[u] = j < nslots ? tab[j] : 0;
            }
        }
        __syncthreads();
        const uint32_t n = pool_n;
        WaveTopK fm;
        if (n > POOL) {
            // every wave offers its share of the table to its own list, then
            // wave 0 merges the four
            fm.init(slots[wid][q], CW, a.k, nullptr, a.k_out);
            for (uint32_t j0 = wid * 64 * 8; j0 < nslots; j0 += 256 * 8) {
                uint64_t kk[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const uint32_t j = j0 + u * 64 + lane;
                    kk[u] = j < nslots ? tab[j] : 0;
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) fm.offer(kk[u], kk[u] != 0 && kk[u] >= floor_key);
            }
            fm.compact();
            if (lane == 0) wave_count[wid][q] = fm.count;
            __syncthreads();
            if (wid == 0) {
                for (uint32_t w = 1; w < 4; ++w) {
                    const uint32_t c = wave_count[w][q];
                    for (uint32_t i0 = 0; i0 < c; i0 += 64) {
                        const uint32_t i = i0 + lane;
                        fm.offer(i < c ? slots[w][q][i] : 0, i < c);
                    }
                }
            }
        } else if (wid == 0) {
            fm.init(pool, POOL, a.k, nullptr, a.k_out);
            fm.count = n;
        }
        if (wid == 0) {
            fm.compact();
            const uint64_t *list = fm.buf;
            uint32_t g = 0, e = 0;
            if ((uint32_t)q < a.nvalid)
                heap_tie_shift(list, fm.count, a.k_out,
                               a.k + 1 >= 2 * a.k_out || a.k >= a.rows, g, e);
            for (uint32_t i = lane; i < a.k_out && (uint32_t)q < a.nvalid; i += 64) {
                const uint32_t src_i = i < g ? i : i + e;
                uint64_t kv = src_i < fm.count ? list[src_i] : 0;
                a.out_idx[(size_t)q * a.out_stride + i] = kv ? key_row(kv) : 0;
                a.out_sim[(size_t)q * a.out_stride + i] = kv ? key_sim(kv) : 0.f;
            }
        }
        // the pool and the slots are reused by the next query
        __syncthreads();
    }
    if (wid == 0 && lane == 0) __hip_atomic_store(a.ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

void launch_to_tiled(const float *src, uint32_t rows, uint32_t dims, uint32_t d4, float4 *dst,
                     uint32_t *flags, hipStream_t s) {
    uint32_t tiles = (rows + 63) / 64;
    hipLaunchKernelGGL(to_tiled_kernel, dim3(tiles), dim3(256), 0, s, src, rows, dims, d4, dst,
                       flags);
    NNDEX_HIP_CHECK(hipGetLastError());
}

void launch_normalize_tiled(float4 *x, uint32_t rows, uint32_t dims, uint32_t d4,
                            uint32_t *flags, hipStream_t s) {
    hipLaunchKernelGGL(normalize_tiled_kernel, dim3((rows + 255) / 256), dim3(256), 0, s, x, rows,
                       dims, d4, flags);
    NNDEX_HIP_CHECK(hipGetLastError());
}

void launch_normalize_rowmajor(const float *src, uint32_t rows, uint32_t dims, uint32_t pdims,
                               int mode, float *dst, uint32_t *flags, uint8_t *row_flags,
                               hipStream_t s) {
    if (rows == 0) return;
    hipLaunchKernelGGL(normalize_rowmajor_kernel, dim3((rows + 255) / 256), dim3(256), 0, s, src,
                       rows, dims, pdims, mode, dst, flags, row_flags);
    NNDEX_HIP_CHECK(hipGetLastError());
}

template <int QB, int CW, bool NT>
static void launch_scan_nt(const ScanArgs &a, uint32_t grid, hipStream_t s) {
    // single-query scans may use larger load groups (more bytes in flight
    // per wave); the group must divide half a row
    const uint32_t code = (a.flags >> kScanGroupShift) & 3u;
    if constexpr (QB == 1) {
        if (code == 2 && a.d4 % 32 == 0) {
            hipLaunchKernelGGL((exact_scan_kernel<QB, CW, 16, NT>), dim3(grid), dim3(256), 0, s, a);
            return;
        }
        if (code == 1 && a.d4 % 24 == 0) {
            hipLaunchKernelGGL((exact_scan_kernel<QB, CW, 12, NT>), dim3(grid), dim3(256), 0, s, a);
            return;
        }
    }
    hipLaunchKernelGGL((exact_scan_kernel<QB, CW, 8, NT>), dim3(grid), dim3(256), 0, s, a);
}

template <int QB, int CW>
static void launch_scan_t(const ScanArgs &a, uint32_t grid, hipStream_t s) {
    if (a.flags & kScanNonTemporal)
        launch_scan_nt<QB, CW, true>(a, grid, s);
    else
        launch_scan_nt<QB, CW, false>(a, grid, s);
    NNDEX_HIP_CHECK(hipGetLastError());
}

__global__ __launch_bounds__(256) void copy_kernel(const uint4 *__restrict__ src,
                                                   uint4 *__restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        dst[i] = src[i];
}

void launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (bytes % 16)
        throw Error(NNDEX_ERR_INVALID_ARGUMENT, "copy size must be a multiple of 16 bytes");
    size_t n = bytes / 16;
    uint32_t grid = (uint32_t)((n + 255) / 256);
    if (grid > 256u * 64) grid = 256u * 64;
    if (grid == 0) return;
    hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, s,
                       static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), n);
    NNDEX_HIP_CHECK(hipGetLastError());
}

uint32_t scan_slots_for_k(uint32_t k) {
    uint32_t cw = 128;
    while (cw < k + 64) cw <<= 1;
    return cw;
}

uint32_t scan_max_qb(uint32_t k) {
    uint32_t cw = scan_slots_for_k(k);
    if (cw > 2048) return 0;
    uint32_t qb = 2048 / cw;  // 4 waves x QB x CW x 8 B <= 64 KiB
    // at most 4 queries per launch: eight run out of scalar registers for
    // their query chunks and take longer than two launches of four
    return qb > 4 ? 4 : qb;
}

void launch_exact_scan(const ScanArgs &a, uint32_t qb, uint32_t grid, hipStream_t s) {
    const uint32_t cw = scan_slots_for_k(a.k);
#define NNDEX_SCAN_CASE(Q, C)                 \
    if (qb == Q && cw == C) {                 \
        launch_scan_t<Q, C>(a, grid, s);      \
        return;                               \
    }
    NNDEX_SCAN_CASE(1, 128)
    NNDEX_SCAN_CASE(1, 256)
    NNDEX_SCAN_CASE(1, 512)
    NNDEX_SCAN_CASE(1, 1024)
    NNDEX_SCAN_CASE(1, 2048)
    NNDEX_SCAN_CASE(2, 128)
    NNDEX_SCAN_CASE(2, 256)
    NNDEX_SCAN_CASE(2, 512)
    NNDEX_SCAN_CASE(2, 1024)
    NNDEX_SCAN_CASE(4, 128)
    NNDEX_SCAN_CASE(4, 256)
    NNDEX_SCAN_CASE(4, 512)
#undef NNDEX_SCAN_CASE
    throw Error(NNDEX_ERR_UNSUPPORTED, "no exact scan variant for qb=" + std::to_string(qb) +
                                           " k=" + std::to_string(a.k));
}

}  // namespace nndex
