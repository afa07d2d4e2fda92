// Front-end `NNdex` (src/lib.rs:160-811) restated in C++ over the HIP backend.
//
// Kept from the reference: shape/k/dims validation and its error order
// (src/lib.rs:185-402, 433-450), padding to a multiple of 8 (:192), k capping
// (:347, :390), and the 64-entry constructor cache keyed by source pointer,
// shape, options and a sampled fingerprint (:601-768).
// Moved to the device: row normalisation of the matrix and of batch queries,
// and single-query preparation (normalize_vector / pad_vector, :404-411,
// :558-599, in the reference's sequential order), with the query checks read
// back together with the results.
// This build carries the GPU backend only, exactly like the reference built
// with `--features gpu` and without `cpu`: BackendPreference::Cpu yields
// NNdexError::CpuBackendUnavailable (src/lib.rs:244-254).
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "gpu_index.hpp"

namespace nndex {
namespace {

constexpr size_t kConstructorCacheCapacity = 64;  // src/lib.rs:602
constexpr float kF32Epsilon = 1.1920929e-7f;

Error shape_mismatch(size_t rows, size_t dims, size_t values) {
    return Error(NNDEX_ERR_SHAPE_MISMATCH, "matrix shape mismatch: rows=" + std::to_string(rows) +
                                               ", dims=" + std::to_string(dims) +
                                               ", values=" + std::to_string(values));
}

// validate_matrix_shape, src/lib.rs:433-442
void validate_matrix_shape(size_t values, size_t rows, size_t dims) {
    unsigned __int128 prod = (unsigned __int128)rows * dims;
    if (rows == 0 || dims == 0 || prod != values) throw shape_mismatch(rows, dims, values);
}

Error gpu_error(const Error &e) {
    // NNdexError::Gpu wraps GpuError with this prefix (src/lib.rs:141-144).
    if (e.code >= NNDEX_ERR_GPU_ADAPTER_NOT_FOUND && e.code <= NNDEX_ERR_GPU_DEVICE_INDEX_OUT_OF_RANGE)
        return Error(e.code, std::string("gpu execution error: ") + e.what());
    return e;
}

// Host restatement of the row checks of normalize_rows (src/lib.rs:519-553),
// used only to raise the reference's error before CpuBackendUnavailable.
void host_validate_rows(const float *m, size_t rows, size_t dims, bool normalized) {
    for (size_t r = 0; r < rows; ++r) {
        const float *s = m + r * dims;
        if (normalized) {
            for (size_t i = 0; i < dims; ++i)
                if (!std::isfinite(s[i])) throw Error(NNDEX_ERR_NON_FINITE_INPUT, "all input values must be finite");
            continue;
        }
        float sq;
        if (dims >= 16) {
            float l[4] = {0, 0, 0, 0};
            size_t i = 0;
            for (; i + 4 <= dims; i += 4)
                for (int j = 0; j < 4; ++j) l[j] = std::fma(s[i + j], s[i + j], l[j]);
            sq = (l[0] + l[1]) + (l[2] + l[3]);
            for (; i < dims; ++i) sq = std::fma(s[i], s[i], sq);
            if (!std::isfinite(sq)) throw Error(NNDEX_ERR_NON_FINITE_INPUT, "all input values must be finite");
        } else {
            sq = 0.f;
            for (size_t i = 0; i < dims; ++i) {
                if (!std::isfinite(s[i])) throw Error(NNDEX_ERR_NON_FINITE_INPUT, "all input values must be finite");
                sq += s[i] * s[i];
            }
        }
        if (sq <= kF32Epsilon) throw Error(NNDEX_ERR_ZERO_NORM, "encountered a zero-norm vector");
    }
}

// sampled_matrix_fingerprint, src/lib.rs:746-768: length, first and last bits.
uint64_t fingerprint(const float *v, size_t n) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](uint64_t x) {
        for (int i = 0; i < 8; ++i) {
            h ^= (x >> (8 * i)) & 0xff;
            h *= 1099511628211ull;
        }
    };
    mix(n);
    if (n) {
        uint32_t a, b;
        std::memcpy(&a, v, 4);
        std::memcpy(&b, v + n - 1, 4);
        mix(a);
        mix(b);
    }
    return h;
}

using CacheKey = std::tuple<uintptr_t, size_t, size_t, bool, bool, uint64_t>;

// BoundedConstructorCache, src/lib.rs:605-667 (LRU, promote on hit).
struct ConstructorCache {
    std::mutex mu;
    std::map<CacheKey, std::shared_ptr<GpuIndex>> entries;
    std::deque<CacheKey> order;

    std::shared_ptr<GpuIndex> get(const CacheKey &k) {
        std::lock_guard<std::mutex> l(mu);
        auto it = entries.find(k);
        if (it == entries.end()) return nullptr;
        promote(k);
        return it->second;
    }
    void insert(const CacheKey &k, std::shared_ptr<GpuIndex> v) {
        std::lock_guard<std::mutex> l(mu);
        auto it = entries.find(k);
        if (it != entries.end()) {
            it->second = std::move(v);
            promote(k);
            return;
        }
        if (entries.size() >= kConstructorCacheCapacity && !order.empty()) {
            entries.erase(order.front());
            order.pop_front();
        }
        entries.emplace(k, std::move(v));
        order.push_back(k);
    }
    size_t size() {
        std::lock_guard<std::mutex> l(mu);
        return entries.size();
    }
    void promote(const CacheKey &k) {
        for (auto it = order.begin(); it != order.end(); ++it)
            if (*it == k) {
                order.erase(it);
                break;
            }
        order.push_back(k);
    }
};

ConstructorCache &gpu_cache() {
    static ConstructorCache c;
    return c;
}

}  // namespace

struct Front {
    size_t rows = 0, dims = 0, pdims = 0;
    bool normalized = false, approx = false;
    std::shared_ptr<GpuIndex> gpu;
};

}  // namespace nndex

using namespace nndex;

struct nndex_index {
    Front f;
    nndex_gpu_index handle;
};

extern "C" {

nndex_status nndex_new(const float *matrix, size_t values, size_t rows, size_t dims,
                       const nndex_index_options *options, nndex_index **out) {
    return guard([&] {
        if (!options || !out || (!matrix && values))
            throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        validate_matrix_shape(values, rows, dims);
        const nndex_index_options o = *options;
        const bool normalized = o.normalized != 0, approx = o.approx != 0;
        const bool enable_cache = o.enable_cache != 0;
        const size_t pdims = (dims + 7) / 8 * 8;
        CacheKey key{reinterpret_cast<uintptr_t>(matrix), rows, dims, normalized, approx,
                     fingerprint(matrix, values)};
        auto make = [&](std::shared_ptr<GpuIndex> g) {
            auto h = new nndex_index();
            h->f.rows = rows;
            h->f.dims = dims;
            h->f.pdims = pdims;
            h->f.normalized = normalized;
            h->f.approx = approx;
            h->f.gpu = g;
            h->handle.impl = std::move(g);
            *out = h;
        };
        if (o.backend == NNDEX_BACKEND_GPU) {
            if (enable_cache) {
                if (auto hit = gpu_cache().get(key)) return make(hit);
            }
            std::shared_ptr<GpuIndex> g;
            try {
                g = GpuIndex::from_raw(matrix, rows, dims, normalized, approx, enable_cache,
                                       o.gpu_device_index);
            } catch (const Error &e) {
                throw gpu_error(e);
            }
            if (enable_cache) gpu_cache().insert(key, g);
            return make(g);
        }
        if (o.backend == NNDEX_BACKEND_CPU) {
            host_validate_rows(matrix, rows, dims, normalized);
            throw Error(NNDEX_ERR_CPU_BACKEND_UNAVAILABLE,
                        "cpu backend is unavailable (crate built without `cpu` feature)");
        }
        throw Error(NNDEX_ERR_INVALID_ARGUMENT, "unknown backend " + std::to_string(o.backend));
    });
}

void nndex_free(nndex_index *index) { delete index; }

int32_t nndex_backend_of(const nndex_index *index) {
    return index && index->f.gpu ? NNDEX_BACKEND_GPU : NNDEX_BACKEND_CPU;
}
size_t nndex_rows(const nndex_index *index) { return index ? index->f.rows : 0; }
size_t nndex_dims(const nndex_index *index) { return index ? index->f.dims : 0; }

nndex_status nndex_gpu_info(const nndex_index *index, nndex_gpu_device_info *out,
                            int32_t *has_info) {
    return guard([&] {
        if (!index || !out || !has_info) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        *has_info = 0;
        if (index->f.gpu) {
            *out = index->f.gpu->info();
            out->index = 0;  // src/lib.rs:306
            *has_info = 1;
        }
    });
}

nndex_status nndex_search(nndex_index *index, const float *query, size_t query_len, size_t k,
                          uint64_t *out_indices, float *out_similarities, size_t *out_len) {
    return guard([&] {
        if (!index || (!query && query_len) || !out_indices || !out_similarities || !out_len)
            throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        const Front &f = index->f;
        if (k == 0) throw Error(NNDEX_ERR_INVALID_TOP_K, "top-k must be greater than zero");
        if (query_len != f.dims)
            throw Error(NNDEX_ERR_QUERY_DIM_MISMATCH,
                        "query dimensionality mismatch: expected " + std::to_string(f.dims) +
                            ", found " + std::to_string(query_len));
        const size_t capped = k < f.rows ? k : f.rows;
        std::vector<Neighbor> res;
        try {
            res = f.gpu->search_raw(query, f.dims, !f.normalized, capped, f.approx);
        } catch (const Error &e) {
            throw gpu_error(e);
        }
        for (size_t i = 0; i < res.size(); ++i) {
            out_indices[i] = res[i].index;
            out_similarities[i] = res[i].similarity;
        }
        *out_len = res.size();
    });
}

nndex_status nndex_search_batch(nndex_index *index, const float *queries, size_t values,
                                size_t query_rows, size_t k, uint64_t *out_indices,
                                float *out_similarities, size_t *out_lens) {
    return guard([&] {
        if (!index || (!queries && values) || !out_indices || !out_similarities || !out_lens)
            throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        const Front &f = index->f;
        validate_matrix_shape(values, query_rows, f.dims);
        if (k == 0) throw Error(NNDEX_ERR_INVALID_TOP_K, "top-k must be greater than zero");
        const size_t capped = k < f.rows ? k : f.rows;
        std::vector<std::vector<Neighbor>> res;
        try {
            res = f.gpu->search_batch_raw(queries, query_rows, f.dims, !f.normalized, capped,
                                          f.approx);
        } catch (const Error &e) {
            throw gpu_error(e);
        }
        for (size_t q = 0; q < res.size(); ++q) {
            for (size_t i = 0; i < res[q].size(); ++i) {
                out_indices[q * capped + i] = res[q][i].index;
                out_similarities[q * capped + i] = res[q][i].similarity;
            }
            out_lens[q] = res[q].size();
        }
    });
}

nndex_gpu_index *nndex_backend_handle(nndex_index *index) {
    return index && index->f.gpu ? &index->handle : nullptr;
}

}  // extern "C"
