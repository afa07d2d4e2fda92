// Shared host/device helpers for the nndex MI355X backend.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "nndex_hip.h"

namespace nndex {

// Error carried from the implementation to the C ABI, where it becomes an
// nndex_status plus a per-thread message.
struct Error : std::runtime_error {
    nndex_status code;
    Error(nndex_status c, const std::string &msg) : std::runtime_error(msg), code(c) {}
};

void set_last_error(const std::string &msg);
void clear_last_error();

#define NNDEX_HIP_CHECK(expr)                                                              \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess) {                                                            \
            throw ::nndex::Error(_e == hipErrorOutOfMemory ? NNDEX_ERR_OUT_OF_MEMORY       \
                                                           : NNDEX_ERR_HIP,                \
                                 std::string(#expr) + ": " + hipGetErrorString(_e));       \
        }                                                                                  \
    } while (0)

// Runs `fn`, converting exceptions into a status and the per-thread message.
template <typename F>
nndex_status guard(F &&fn) {
    try {
        clear_last_error();
        fn();
        return NNDEX_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return NNDEX_ERR_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return NNDEX_ERR_HIP;
    }
}

// Native 4-wide float vector (HIP's float4 is a struct wrapper around it).
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Streaming 16-byte load that bypasses cache retention (data read once).
__device__ inline float4 load_nt(const float4 *p) {
    f32x4 v = __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}

// Read-only view in the constant address space: wave-uniform loads through it
// become scalar (s_load) loads even when the kernel also writes memory.
template <typename T>
using const_as = const __attribute__((address_space(4))) T;
template <typename T>
__device__ inline const_as<T> *as_constant(const T *p) {
    return (const_as<T> *)(p);
}

// ---- 64-bit ranking keys ---------------------------------------------------
// key = ordered(similarity) << 32 | (0xFFFFFFFF - row).  Unsigned key order is
// (similarity by f32::total_cmp, then lower row first), i.e. the output order
// of src/topk.rs:143-148.  Key 0 never encodes a finite score and marks "empty".
__host__ __device__ inline uint32_t ordered_bits(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float from_ordered(uint32_t o) {
    uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    return __builtin_bit_cast(float, u);
}
__host__ __device__ inline uint64_t make_key(float sim, uint32_t row) {
    return (static_cast<uint64_t>(ordered_bits(sim)) << 32) | (0xFFFFFFFFu - row);
}
__host__ __device__ inline uint32_t key_row(uint64_t key) {
    return 0xFFFFFFFFu - static_cast<uint32_t>(key);
}
__host__ __device__ inline float key_sim(uint64_t key) {
    return from_ordered(static_cast<uint32_t>(key >> 32));
}
// A float floor f with: make_key(s, row) >= threshold implies !(s < f), NaN
// scores included (the test admits a superset of the passing keys): -inf for
// thresholds at or below -inf's keys, +inf for thresholds at or above +inf's
// keys (only +inf and NaN scores get through), else the threshold's own
// similarity.
__host__ __device__ inline float key_floor_sim(uint64_t threshold) {
    const uint32_t hi = static_cast<uint32_t>(threshold >> 32);
    if (hi <= ordered_bits(-__builtin_inff())) return -__builtin_inff();
    if (hi >= ordered_bits(__builtin_inff())) return __builtin_inff();
    return from_ordered(hi);
}

// Rows are stored in tiles of 64; within a tile, float4 chunk c of row r sits
// at ((tile * d4 + c) * 64 + r): a wavefront reading chunk c of its 64 rows
// reads 1 KiB contiguously, and every lane walks its own row in order.
constexpr uint32_t kTileRows = 64;
__host__ __device__ inline size_t tiled_offset4(size_t row, size_t c4, size_t d4) {
    return ((row / kTileRows) * d4 + c4) * kTileRows + (row % kTileRows);
}
inline size_t padded_rows(size_t rows) { return (rows + kTileRows - 1) / kTileRows * kTileRows; }

}  // namespace nndex
