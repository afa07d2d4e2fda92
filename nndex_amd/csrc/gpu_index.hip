// GpuIndex implementation and the backend C ABI (replaces src/gpu.rs).
#include "gpu_index.hpp"

#include <cmath>
#include <cstring>
#include <string>

namespace nndex {

// ---------------------------------------------------------------------------
// Per-thread error message
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }
void clear_last_error() { g_last_error.clear(); }

// Result caches of src/gpu.rs:21-25: cleared wholesale when full.
constexpr size_t kSingleCacheCapacity = 128;
constexpr size_t kBatchCacheCapacity = 16;
// Bytes per host -> device staging chunk during construction.
constexpr size_t kUploadChunkBytes = size_t(16) << 20;

void DeviceBuffer::reserve(size_t n) {
    if (n <= bytes) return;
    if (ptr) NNDEX_HIP_CHECK(hipFree(ptr));
    ptr = nullptr;
    bytes = 0;
    NNDEX_HIP_CHECK(hipMalloc(&ptr, n));
    bytes = n;
}
DeviceBuffer::~DeviceBuffer() {
    if (ptr) (void)hipFree(ptr);
}

void PinnedBuffer::reserve(size_t n) {
    if (n <= bytes) return;
    if (ptr) NNDEX_HIP_CHECK(hipHostFree(ptr));
    ptr = nullptr;
    bytes = 0;
    NNDEX_HIP_CHECK(hipHostMalloc(&ptr, n, hipHostMallocDefault));
    bytes = n;
}
PinnedBuffer::~PinnedBuffer() {
    if (ptr) (void)hipHostFree(ptr);
}

template <typename T>
static void regrow(T *&p, size_t n) {
    if (p) NNDEX_HIP_CHECK(hipFree(p));
    p = nullptr;
    NNDEX_HIP_CHECK(hipMalloc(&p, n * sizeof(T) + 256));
}

Workspace::~Workspace() {
    for (void *p : {(void *)part, (void *)part_count, (void *)theta, (void *)ticket, (void *)qlow,
                    (void *)qimg, (void *)thresh_f, (void *)delta, (void *)qscale,
                    (void *)scores, (void *)thresh, (void *)cand, (void *)cand_count,
                    (void *)flags, (void *)ivf_probe, (void *)ivf_probe_sim, (void *)ivf_part,
                    (void *)xflags, (void *)ivf_pair_u32, (void *)ivf_list_u32, ivf_sort_tmp,
                    (void *)sel_keys, (void *)sel_cand, (void *)sel_sorted, (void *)sel_st,
                    (void *)sel_hist, (void *)sel_cnt, (void *)sel_offs, (void *)sel_n,
                    sel_tmp})
        if (p) (void)hipFree(p);
}

void Workspace::reserve_select(size_t nq, size_t ld, size_t ksel) {
    // every buffer sized for the largest group, stride and k seen so far
    // (the layout strides are sel_ld and sel_k whatever the call asks)
    const size_t nn = nq > sel_nq ? nq : sel_nq, nld = ld > sel_ld ? ld : sel_ld,
                 nk = ksel > sel_k ? ksel : sel_k;
    if (nn == sel_nq && nld == sel_ld && nk == sel_k) return;
    ++generation;
    if (nn != sel_nq || nld != sel_ld) regrow(sel_keys, nn * nld);
    if (nn != sel_nq || nk != sel_k) {
        regrow(sel_cand, nn * nk);
        regrow(sel_sorted, nn * nk);
    }
    if (nn != sel_nq) {
        // histograms are cleared by each selection's first kernel
        regrow(sel_st, nn);
        regrow(sel_hist, nn * kSelectBins);
        regrow(sel_cnt, nn);
        regrow(sel_offs, 2 * nn);
        regrow(sel_n, nn);
    }
    sel_nq = nn;
    sel_ld = nld;
    sel_k = nk;
    const size_t bytes = select_sort_bytes(sel_nq, sel_k);
    if (bytes > sel_tmp_bytes) {
        if (sel_tmp) NNDEX_HIP_CHECK(hipFree(sel_tmp));
        sel_tmp = nullptr;
        NNDEX_HIP_CHECK(hipMalloc(&sel_tmp, bytes));
        sel_tmp_bytes = bytes;
    }
}

SelectBuffers Workspace::select_buffers(uint32_t nq, uint32_t ksel, const uint32_t *n,
                                        uint32_t n_all) const {
    if (nq > sel_nq || ksel > sel_k)
        throw Error(NNDEX_ERR_INVALID_ARGUMENT, "selection buffers are smaller than the launch");
    SelectBuffers b;
    b.keys = sel_keys;
    b.ld = sel_ld;
    b.n = n;
    b.n_all = n_all;
    b.nq = nq;
    b.ksel = ksel;
    b.st = sel_st;
    b.hist = sel_hist;
    b.cand = sel_cand;
    b.sorted = sel_sorted;
    b.cnt = sel_cnt;
    b.offs = sel_offs;
    b.sort_tmp = sel_tmp;
    b.sort_bytes = sel_tmp_bytes;
    return b;
}

void Workspace::reserve_batch(size_t nq, size_t pdims, size_t rows_s, size_t capq) {
    rows_s = sample_ld(rows_s);
    if (nq > b_queries || nq * rows_s > b_scores || nq * capq > b_cand) ++generation;
    if (nq > b_queries) {
        regrow(qlow, nq * pdims / 8);
        regrow(qimg, mfma_image_units(nq));
        regrow(delta, nq);
        // thresholds and FP8 scales are copied a whole query chunk (<= 64)
        // at a time by the MFMA scan
        regrow(thresh, nq + 64);
        regrow(thresh_f, nq + 64);
        regrow(qscale, nq + 64);
        regrow(cand_count, nq);
        regrow(flags, nq);
        b_queries = nq;
        b_scores = 0;
        b_cand = 0;
    }
    if (nq * rows_s > b_scores) {
        regrow(scores, nq * rows_s);
        b_scores = nq * rows_s;
    }
    if (nq * capq > b_cand) {
        regrow(cand, nq * capq);
        b_cand = nq * capq;
    }
}

void Workspace::reserve_ivf(size_t pairs, size_t nlist, size_t part_elems, size_t sort_bytes) {
    if (pairs > ivf_pairs) {
        regrow(ivf_probe, pairs);
        regrow(ivf_probe_sim, pairs);
        regrow(ivf_pair_u32, 5 * pairs);
        ivf_pairs = pairs;
    }
    if (nlist > ivf_lists) {
        regrow(ivf_list_u32, 4 * (nlist + 1));
        ivf_lists = nlist;
    }
    if (part_elems > ivf_part_elems) {
        regrow(ivf_part, part_elems);
        ivf_part_elems = part_elems;
    }
    if (sort_bytes > ivf_sort_bytes) {
        if (ivf_sort_tmp) NNDEX_HIP_CHECK(hipFree(ivf_sort_tmp));
        ivf_sort_tmp = nullptr;
        NNDEX_HIP_CHECK(hipMalloc(&ivf_sort_tmp, sort_bytes));
        ivf_sort_bytes = sort_bytes;
    }
}

int device_count() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

void fill_device_info(int device, nndex_gpu_device_info &out) {
    std::memset(&out, 0, sizeof(out));
    hipDeviceProp_t p;
    NNDEX_HIP_CHECK(hipGetDeviceProperties(&p, device));
    int rt = 0, drv = 0;
    (void)hipRuntimeGetVersion(&rt);
    (void)hipDriverGetVersion(&drv);
    out.index = static_cast<uint64_t>(device);
    if (p.name[0])
        std::snprintf(out.name, sizeof(out.name), "%s", p.name);
    else  // some drivers report an empty marketing name
        std::snprintf(out.name, sizeof(out.name), "AMD Instinct GPU (%.40s)", p.gcnArchName);
    std::snprintf(out.driver, sizeof(out.driver), "HIP %d", drv);
    std::snprintf(out.driver_info, sizeof(out.driver_info), "runtime %d, %s, %d CUs", rt,
                  p.gcnArchName, p.multiProcessorCount);
    std::snprintf(out.backend, sizeof(out.backend), "Hip");
    std::snprintf(out.device_type, sizeof(out.device_type), "%s",
                  p.integrated ? "IntegratedGpu" : "DiscreteGpu");
}

void throw_flags(uint32_t flags) {
    if (flags & 1u) throw Error(NNDEX_ERR_NON_FINITE_INPUT, "all input values must be finite");
    if (flags & 2u) throw Error(NNDEX_ERR_ZERO_NORM, "encountered a zero-norm vector");
}

// select_adapter (src/gpu.rs:933-970) for HIP devices.
void GpuIndex::init_device(int64_t device_index) {
    int count = device_count();
    if (device_index >= 0) {
        if (device_index >= count)
            throw Error(NNDEX_ERR_GPU_DEVICE_INDEX_OUT_OF_RANGE,
                        "gpu device index " + std::to_string(device_index) +
                            " is out of range (found " + std::to_string(count) + " adapter(s))");
        device_ = static_cast<int>(device_index);
    } else {
        if (count == 0)
            throw Error(NNDEX_ERR_GPU_ADAPTER_NOT_FOUND, "no compatible GPU adapter was found");
        device_ = 0;
    }
    NNDEX_HIP_CHECK(hipSetDevice(device_));
    fill_device_info(device_, info_);
    hipDeviceProp_t p;
    NNDEX_HIP_CHECK(hipGetDeviceProperties(&p, device_));
    ncu_ = static_cast<uint32_t>(p.multiProcessorCount);
    NNDEX_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
}

GpuIndex::~GpuIndex() {
    if (stream_) {
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamDestroy(stream_);
    }
    if (x_) (void)hipFree(x_);
    if (xbf16_) (void)hipFree(xbf16_);
    if (xrow_) (void)hipFree(xrow_);
    if (xmax_sq_) (void)hipFree(xmax_sq_);
    if (xerr_sq_) (void)hipFree(xerr_sq_);
    for (void *p : {(void *)xfp8_, (void *)xscale_, (void *)xerr8_sq_})
        if (p) (void)hipFree(p);
    for (auto &smp : samples_) {
        if (smp.f32) (void)hipFree(smp.f32);
        if (smp.bf16) (void)hipFree(smp.bf16);
        for (void *p : {(void *)smp.fp8, (void *)smp.fp8_scale, (void *)smp.fp8_err})
            if (p) (void)hipFree(p);
    }
}

void GpuIndex::upload_tiled(const float *host, size_t rows, size_t src_dims, bool check_finite,
                            bool normalize, size_t logical_dims) {
    const size_t tiles = padded_rows(rows) / kTileRows;
    NNDEX_HIP_CHECK(hipMalloc(&x_, tiles * d4_ * kTileRows * sizeof(float4)));
    // Rows go host -> pinned staging -> device in chunks of about 16 MiB (a
    // multiple of 64 rows): the host copy of chunk i + 1 (from a numpy array
    // or a memory-mapped file) overlaps the DMA and tiling of chunk i, two
    // pinned and two device buffers in turn, an event guarding each reuse.
    size_t chunk = (kUploadChunkBytes / (src_dims * sizeof(float))) / kTileRows * kTileRows;
    if (chunk < kTileRows) chunk = kTileRows;
    if (chunk > rows) chunk = rows;
    DeviceBuffer staging[2], flags;
    PinnedBuffer pinned[2];
    hipEvent_t done[2] = {nullptr, nullptr};
    auto release = [&] {
        for (hipEvent_t e : done)
            if (e) (void)hipEventDestroy(e);
    };
    try {
        for (int b = 0; b < 2; ++b) {
            staging[b].reserve(chunk * src_dims * sizeof(float));
            pinned[b].reserve(chunk * src_dims * sizeof(float));
            NNDEX_HIP_CHECK(hipEventCreateWithFlags(&done[b], hipEventDisableTiming));
        }
        flags.reserve(sizeof(uint32_t));
        launch_zero_u32(static_cast<uint32_t *>(flags.ptr), 1, stream_);
        size_t i = 0;
        for (size_t r0 = 0; r0 < rows; r0 += chunk, ++i) {
            const int b = static_cast<int>(i & 1);
            const size_t n = rows - r0 < chunk ? rows - r0 : chunk;
            const size_t bytes = n * src_dims * sizeof(float);
            if (i >= 2) NNDEX_HIP_CHECK(hipEventSynchronize(done[b]));  // chunk i - 2 is done
            std::memcpy(pinned[b].ptr, host + r0 * src_dims, bytes);
            NNDEX_HIP_CHECK(hipMemcpyAsync(staging[b].ptr, pinned[b].ptr, bytes,
                                           hipMemcpyHostToDevice, stream_));
            launch_to_tiled(static_cast<const float *>(staging[b].ptr), static_cast<uint32_t>(n),
                            static_cast<uint32_t>(src_dims), static_cast<uint32_t>(d4_),
                            x_ + (r0 / kTileRows) * d4_ * kTileRows,
                            static_cast<uint32_t *>(flags.ptr), stream_);
            NNDEX_HIP_CHECK(hipEventRecord(done[b], stream_));
        }
        if (normalize)
            launch_normalize_tiled(x_, static_cast<uint32_t>(rows),
                                   static_cast<uint32_t>(logical_dims), static_cast<uint32_t>(d4_),
                                   static_cast<uint32_t *>(flags.ptr), stream_);
        NNDEX_HIP_CHECK(hipMalloc(&xmax_sq_, sizeof(unsigned long long)));
        launch_zero_u32(reinterpret_cast<uint32_t *>(xmax_sq_), 2, stream_);
        launch_max_norm(x_, static_cast<uint32_t>(rows), static_cast<uint32_t>(d4_), xmax_sq_,
                        stream_);
        uint32_t *h_flags = static_cast<uint32_t *>(pinned[0].ptr);
        NNDEX_HIP_CHECK(hipMemcpyAsync(h_flags, flags.ptr, sizeof(uint32_t),
                                       hipMemcpyDeviceToHost, stream_));
        NNDEX_HIP_CHECK(hipStreamSynchronize(stream_));
        release();
        if (check_finite || normalize) throw_flags(*h_flags);
    } catch (...) {
        (void)hipStreamSynchronize(stream_);
        release();
        throw;
    }
}

void GpuIndex::ensure_bf16(hipStream_t s) {
    std::lock_guard<std::mutex> l(lazy_mu_);
    if (xbf16_) return;
    const size_t units = padded_rows(rows_) * (pdims_ / 8);
    NNDEX_HIP_CHECK(hipMalloc(&xbf16_, units * sizeof(uint4)));
    launch_to_bf16(x_, padded_rows(rows_), static_cast<uint32_t>(pdims_ / 8), xbf16_, s);
    NNDEX_HIP_CHECK(hipMalloc(&xerr_sq_, 2 * sizeof(unsigned long long)));
    NNDEX_HIP_CHECK(hipMemsetAsync(xerr_sq_, 0, 2 * sizeof(unsigned long long), s));
    launch_bf16_error(x_, static_cast<uint32_t>(rows_), static_cast<uint32_t>(d4_), xerr_sq_ + 1,
                      s);
    // complete before the lock is released: another stream may use it next
    NNDEX_HIP_CHECK(hipStreamSynchronize(s));
}

void GpuIndex::ensure_rowmajor(hipStream_t s) {
    std::lock_guard<std::mutex> l(lazy_mu_);
    if (xrow_) return;
    NNDEX_HIP_CHECK(hipMalloc(&xrow_, rows_ * d4_ * sizeof(float4)));
    launch_tiled_to_rows(x_, static_cast<uint32_t>(rows_), static_cast<uint32_t>(d4_), xrow_, s);
    // complete before the lock is released: another stream may use it next
    NNDEX_HIP_CHECK(hipStreamSynchronize(s));
}

void GpuIndex::ensure_fp8(hipStream_t s) {
    std::lock_guard<std::mutex> l(lazy_mu_);
    if (xfp8_) return;
    const size_t prows = padded_rows(rows_), d16 = (d4_ + 3) / 4;
    NNDEX_HIP_CHECK(hipMalloc(&xfp8_, prows * d16 * sizeof(uint4)));
    NNDEX_HIP_CHECK(hipMalloc(&xscale_, prows * sizeof(float)));
    NNDEX_HIP_CHECK(hipMalloc(&xerr8_sq_, sizeof(unsigned long long)));
    NNDEX_HIP_CHECK(hipMemsetAsync(xerr8_sq_, 0, sizeof(unsigned long long), s));
    launch_to_fp8(x_, prows, static_cast<uint32_t>(d4_), xfp8_, xscale_, xerr8_sq_, s);
    // complete before the lock is released: another stream may use it next
    NNDEX_HIP_CHECK(hipStreamSynchronize(s));
}

const RowSample &GpuIndex::sample_for(uint32_t stride, Precision p, hipStream_t s) {
    std::lock_guard<std::mutex> l(lazy_mu_);
    RowSample *hit = nullptr;
    for (auto &smp : samples_)
        if (smp.stride == stride) hit = &smp;
    if (!hit) {
        samples_.push_back(RowSample{});
        hit = &samples_.back();
        hit->stride = stride;
        hit->rows = static_cast<uint32_t>((rows_ + stride - 1) / stride);
        const size_t prows = padded_rows(hit->rows);
        NNDEX_HIP_CHECK(hipMalloc(&hit->f32, prows * d4_ * sizeof(float4)));
        launch_gather_rows(x_, static_cast<uint32_t>(d4_), stride, hit->rows, hit->f32, s);
        NNDEX_HIP_CHECK(hipStreamSynchronize(s));
    }
    if (p == Precision::FP8 && !hit->fp8) {
        const size_t prows = padded_rows(hit->rows), d16 = (d4_ + 3) / 4;
        NNDEX_HIP_CHECK(hipMalloc(&hit->fp8, prows * d16 * sizeof(uint4)));
        NNDEX_HIP_CHECK(hipMalloc(&hit->fp8_scale, prows * sizeof(float)));
        NNDEX_HIP_CHECK(hipMalloc(&hit->fp8_err, sizeof(unsigned long long)));
        NNDEX_HIP_CHECK(hipMemsetAsync(hit->fp8_err, 0, sizeof(unsigned long long), s));
        launch_to_fp8(hit->f32, prows, static_cast<uint32_t>(d4_), hit->fp8, hit->fp8_scale,
                      hit->fp8_err, s);
        NNDEX_HIP_CHECK(hipStreamSynchronize(s));
    }
    if (p == Precision::BF16 && !hit->bf16) {
        const size_t prows = padded_rows(hit->rows);
        NNDEX_HIP_CHECK(hipMalloc(&hit->bf16, prows * (pdims_ / 8) * sizeof(uint4)));
        launch_to_bf16(hit->f32, prows, static_cast<uint32_t>(pdims_ / 8), hit->bf16, s);
        NNDEX_HIP_CHECK(hipStreamSynchronize(s));
    }
    return *hit;
}

// 16-byte units per row of the MFMA scan's element type.
static size_t mfma_units(size_t pdims, Precision p) {
    return p == Precision::F32 ? pdims / 4 : (p == Precision::BF16 ? pdims / 8 : (pdims + 15) / 16);
}

bool GpuIndex::mfma_supported(Precision p) const { return (mfma_units(pdims_, p) + 3) / 4 <= 32; }

size_t GpuIndex::default_overfetch(size_t k, Precision p) {
    size_t kp = p == Precision::F32 ? k + (k / 2 > 16 ? k / 2 : 16)
                                    : (4 * k > k + 64 ? 4 * k : k + 64);
    return kp > 447 ? 447 : kp;
}

void GpuIndex::search_batch_mfma(Workspace &ws, const float *dq, size_t nq, size_t k, Precision p,
                                 size_t kp, uint64_t *d_idx, float *d_sim, uint8_t *d_flags,
                                 hipStream_t s) {
    const bool bf16 = p == Precision::BF16;
    if (!mfma_supported(p))
        throw Error(NNDEX_ERR_UNSUPPORTED, "dims too large for the MFMA batch path");
    if (k == 0 || k > rows_) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "k must be in [1, rows]");
    const bool fp8 = p == Precision::FP8;
    const int elem = static_cast<int>(p);
    const BatchPlan plan = plan_batch(k, p, kp);
    kp = plan.kp;
    if (bf16) ensure_bf16(s);
    if (fp8) ensure_fp8(s);
    ensure_rowmajor(s);
    const RowSample &smp = sample_for(plan.stride, p, s);
    const size_t capq = plan.capq;
    const size_t group = plan.group;
    const uint32_t dch = static_cast<uint32_t>(mfma_units(pdims_, p));
    for (size_t q0 = 0; q0 < nq; q0 += group) {
        const size_t n = nq - q0 < group ? nq - q0 : group;
        ws.reserve_batch(n, pdims_, smp.rows, capq);
        const float *qf = dq + q0 * pdims_;
        launch_prep_queries(qf, static_cast<uint32_t>(n), static_cast<uint32_t>(pdims_), elem,
                            ws.qlow, ws.qscale, xmax_sq_,
                            bf16 ? xerr_sq_ + 1 : (fp8 ? xerr8_sq_ : xmax_sq_), ws.delta, s);
        const uint4 *qu = p != Precision::F32 ? ws.qlow : reinterpret_cast<const uint4 *>(qf);
        const int variant = mfma_variant_[elem].load(std::memory_order_relaxed);
        // the chunk image both scans copy their query chunks from (its
        // padding thresholds never pass; padding scales are zero)
        uint32_t qc = 0, ru = 0;
        mfma_chunk_shape(elem, dch, variant, qc, ru);
        launch_chunk_image(qu, dch, static_cast<uint32_t>(n), dch, qc, ru, ws.qimg, ws.thresh_f,
                           fp8 ? ws.qscale : nullptr, s);
        // stage 1: sample scores and thresholds
        MfmaScanArgs a{};
        a.qimg = ws.qimg;
        a.x = bf16 ? smp.bf16 : (fp8 ? smp.fp8 : reinterpret_cast<const uint4 *>(smp.f32));
        a.row_scale = smp.fp8_scale;
        a.q_scale = ws.qscale;
        a.rows = smp.rows;
        a.rows_padded = static_cast<uint32_t>(padded_rows(smp.rows));
        a.dch = dch;
        a.q = qu;
        a.q_stride = dch;
        a.nq = static_cast<uint32_t>(n);
        a.mode = 1;
        a.out = ws.scores;
        a.ldo = static_cast<uint32_t>(sample_ld(smp.rows));
        const uint32_t br = mfma_block_rows(elem, dch, false, variant);
        // A large sample writes only each group's largest score (32 groups
        // of br / 32 rows per block tile): the jt-th best group maximum is
        // still a score that at least jt sample rows reach (jt groups, one
        // row each), and almost always the jt-th best row itself, since the
        // best sample rows seldom share a group; the scan writes and the
        // threshold reads br / 32 times fewer scores.
        uint32_t n_sample = smp.rows;
        const uint32_t n_groups = static_cast<uint32_t>((a.rows_padded + br - 1) / br) * 32;
        if (smp.rows / (br / 32) >= 16 * plan.jt && n_groups <= a.ldo) {
            a.mode = 3;
            n_sample = n_groups;
        }
        // one 8-wave block per CU (register-bound); a scan with fewer row
        // tiles than CUs also splits the queries over blocks (grid y),
        // choosing the split that finishes in the fewest block waves per
        // unit of work; with more tiles, the tiles of the last partial round
        // are split the same way (tail_split), so no block idles a round
        const uint32_t nch = static_cast<uint32_t>((n + qc - 1) / qc);
        auto grid_for = [&](uint32_t rp) {
            const uint32_t blocks = (rp + br - 1) / br;
            const uint32_t cap = ncu_;
            a.qsplit = 1;
            a.tail_split = 1;
            if (blocks > cap && blocks % cap) {
                // rounds of tile work, each piece also paying its own loads
                // of the stationary rows (about 2% of a tile)
                const uint32_t full = blocks / cap, rem = blocks % cap;
                double best = full + 1.0;
                for (uint32_t q = 2; q <= 8 && q <= nch; ++q) {
                    const double t = full + (double)((rem * q + cap - 1) / cap) / q + 0.02 * q;
                    if (t < best - 1e-9) {
                        best = t;
                        a.tail_split = q;
                    }
                }
            }
            if (blocks < cap) {
                const uint32_t max_split = static_cast<uint32_t>((n + 15) / 16);
                double best = 1.0;
                for (uint32_t q = 2; q <= 8 && q <= max_split; ++q) {
                    const double t = (double)((blocks * q + cap - 1) / cap) / q;
                    if (t < best - 1e-9) {
                        best = t;
                        a.qsplit = q;
                    }
                }
            }
            return blocks < cap ? blocks : cap;
        };
        launch_mfma_scan(a, elem, grid_for(a.rows_padded), s, variant);
        launch_sample_threshold(ws.scores, static_cast<uint32_t>(n), n_sample, a.ldo,
                                static_cast<uint32_t>(plan.jt), ws.delta, ws.thresh, ws.thresh_f, s);
        // stage 2: filter scan over every row
        launch_zero_u32(ws.cand_count, n, s);
        a.x = bf16 ? xbf16_ : (fp8 ? xfp8_ : reinterpret_cast<const uint4 *>(x_));
        a.row_scale = xscale_;
        a.rows = static_cast<uint32_t>(rows_);
        a.rows_padded = static_cast<uint32_t>(padded_rows(rows_));
        a.mode = 0;
        a.thresh = ws.thresh;
        a.thresh_f = ws.thresh_f;
        a.cand = ws.cand;
        a.cand_count = ws.cand_count;
        a.capq = static_cast<uint32_t>(capq);
        launch_mfma_scan(a, elem, grid_for(a.rows_padded), s, variant);
        // stages 3-4: exact rerank and certificate
        RerankArgs r{};
        r.x = xrow_;
        r.d4 = static_cast<uint32_t>(d4_);
        r.q = reinterpret_cast<const float4 *>(qf);
        r.cand = ws.cand;
        r.cand_count = ws.cand_count;
        r.capq = static_cast<uint32_t>(capq);
        r.thresh = ws.thresh;
        r.delta = ws.delta;
        r.k = static_cast<uint32_t>(k);
        r.kp = static_cast<uint32_t>(kp);
        r.out_idx = d_idx + q0 * k;
        r.out_sim = d_sim + q0 * k;
        r.flags = d_flags + q0;
        launch_rerank(r, static_cast<uint32_t>(n), s);
        // FP8's error bound rarely certifies: its result is the exact rerank
        // of the over-fetched candidates, and the flag reports it unproven
        if (!fp8) launch_rerank_full(r, static_cast<uint32_t>(n), s);
    }
}

BatchGraph::~BatchGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
}

void BatchGraph::launch(hipStream_t s) const { NNDEX_HIP_CHECK(hipGraphLaunch(exec, s)); }

std::unique_ptr<BatchGraph> GpuIndex::capture_batch(Workspace &ws, const float *dq, size_t nq,
                                                    size_t k, Precision p, size_t kp,
                                                    uint64_t *d_idx, float *d_sim,
                                                    uint8_t *d_flags, hipStream_t s) {
    if (!s) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "capture needs a non-null stream");
    NNDEX_HIP_CHECK(hipSetDevice(device_));
    search_batch_mfma(ws, dq, nq, k, p, kp, d_idx, d_sim, d_flags, s);
    NNDEX_HIP_CHECK(hipStreamSynchronize(s));
    std::unique_ptr<BatchGraph> g(new BatchGraph());
    NNDEX_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        search_batch_mfma(ws, dq, nq, k, p, kp, d_idx, d_sim, d_flags, s);
    } catch (...) {
        hipGraph_t dead = nullptr;
        (void)hipStreamEndCapture(s, &dead);
        if (dead) (void)hipGraphDestroy(dead);
        throw;
    }
    NNDEX_HIP_CHECK(hipStreamEndCapture(s, &g->graph));
    NNDEX_HIP_CHECK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    g->generation = ws.generation;
    return g;
}

BatchPlan GpuIndex::plan_batch(size_t k, Precision p, size_t kp) const {
    BatchPlan b;
    if (kp == 0) kp = default_overfetch(k, p);
    if (kp < k) kp = k;
    if (kp > rows_) kp = rows_;
    if (kp > 447) throw Error(NNDEX_ERR_UNSUPPORTED, "over-fetch above 447 is not supported");
    b.kp = kp;
    // sample: about max(128 k', rows/32) rows; the threshold's rank in it
    // (below) sets how many rows pass the filter
    const size_t target = 128 * kp > rows_ / 32 ? 128 * kp : rows_ / 32;
    size_t stride = rows_ / (target ? target : 1);
    if (stride < 1) stride = 1;
    if (stride > 64) stride = 64;
    b.stride = static_cast<uint32_t>(stride);
    b.rows_s = (rows_ + stride - 1) / stride;
    size_t jt = (2 * (kp + 1) + stride - 1) / stride;
    if (jt < 8) jt = 8;
    b.jt = jt < kp ? jt : kp;
    b.capq = 8 * kp * stride;
    if (b.capq < 8192) b.capq = 8192;
    return b;
}

size_t BatchPlan::workspace_bytes(size_t nq, size_t pdims) const {
    const size_t n = nq < group ? nq : group;
    return n * (pdims / 8 * 16 + 4 + 8 + 4 + 4 + 1) + (n + 64) * (8 + 4 + 4) +
           mfma_image_units(n) * 16 + n * sample_ld(rows_s) * 4 + n * capq * 8;
}

bool GpuIndex::exact_batch_precision(Precision &p) const {
    // BF16 scoring for large indexes: its certificate proves most queries'
    // results equal the FP32 search (the rest take the fused scan), at a
    // quarter of the FP32 MFMA time and half the bytes.  Small indexes (the
    // IVF centroid index among them) keep FP32, whose tighter bound leaves
    // fewer queries to the scan where the scoring itself is cheap.
    const bool f32 = mfma_supported(Precision::F32);
    if (mfma_supported(Precision::BF16) && (rows_ >= kBf16ExactMinRows || !f32)) {
        p = Precision::BF16;
        return true;
    }
    if (f32) {
        p = Precision::F32;
        return true;
    }
    return false;
}

void Workspace::reserve_xflags(size_t nq) {
    if (nq <= xflags_n) return;
    ++generation;
    if (xflags) NNDEX_HIP_CHECK(hipFree(xflags));
    xflags = nullptr;
    NNDEX_HIP_CHECK(hipMalloc(&xflags, nq + 256));
    xflags_n = nq;
}

void GpuIndex::prepare(Workspace &ws, size_t nq, size_t k, size_t ivf_nprobe, hipStream_t s) {
    NNDEX_HIP_CHECK(hipSetDevice(device_));
    if (k > ws.max_k) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "k exceeds the workspace's max_k");
    if (k > rows_) k = rows_;
    // large k: the selection buffers for the largest k (each k up to it
    // selects fewer keys)
    for (size_t kk : {k, k - 1})
        if (kk >= 1 && !fused_scan_serves(kk)) {
            ws.reserve_select(4, rows_, 2 * kk - 1 < rows_ ? 2 * kk - 1 : rows_);
            break;
        }
    Precision p;
    if (nq > kFusedBatchMax && exact_batch_precision(p)) {
        // every k up to the maximum: the sample stride and the candidate
        // capacity are not monotone in k
        const bool bf16 = p == Precision::BF16;
        if (bf16) ensure_bf16(s);
        ensure_rowmajor(s);
        size_t rows_s = 0, capq = 0, group = 0;
        for (size_t kk = 1; kk <= k && kk <= 400 && kk <= rows_; ++kk) {
            const BatchPlan plan = plan_batch(kk, p, 0);
            sample_for(plan.stride, p, s);
            rows_s = plan.rows_s > rows_s ? plan.rows_s : rows_s;
            capq = plan.capq > capq ? plan.capq : capq;
            group = plan.group;
        }
        if (group) ws.reserve_batch(nq < group ? nq : group, pdims_, rows_s, capq);
        ws.reserve_xflags(nq);
    }
    if (ivf_nprobe) {
        const IvfIndex &ivf = build_ivf();
        if (ivf.nlist >= 2) {
            const size_t np = ivf_nprobe < ivf.nlist ? ivf_nprobe : ivf.nlist;
            // every (batch size, nprobe, k) up to the maximum
            size_t part = 0;
            for (size_t n = 1; n <= nq; n = n < 16 ? n + 1 : (2 * n < nq || n == nq ? 2 * n : nq))
                for (size_t p = 1; p <= np; ++p)
                    for (size_t kk : {(size_t)1, (size_t)10, k}) {
                        if (kk > k) continue;
                        uint32_t qb, nchunk;
                        const size_t ks = ivf_select_k(kk);
                        ivf_scan_shape(n, p, ivf.nlist, rows_, ks, qb, nchunk);
                        const size_t e = n * p * nchunk * ks;
                        part = e > part ? e : part;
                    }
            ws.reserve_ivf(nq * np, ivf.nlist, part, ivf_sort_bytes(nq * np, ivf.nlist));
            if (ivf_large_k(k)) {
                // the largest group (fewest probes) and the longest key arrays
                size_t gmax = 0, ldmax = 0;
                for (size_t p = 1; p <= np; ++p) {
                    size_t g = 0, ld = 0;
                    ivf_large_k_shape(ivf, nq, p, k, g, ld);
                    gmax = g > gmax ? g : gmax;
                    ldmax = ld > ldmax ? ld : ldmax;
                }
                ws.reserve_select(gmax, ldmax, 2 * k - 1);
            }
            if (!ws.sub || ws.sub->max_k < np)
                ws.sub = ivf.cindex->make_workspace(8, np < 10 ? 10 : np);
            ivf.cindex->prepare(*ws.sub, nq, np, 0, s);
        }
    }
    NNDEX_HIP_CHECK(hipStreamSynchronize(s));
}

size_t GpuIndex::workspace_bytes(size_t nq, size_t k, size_t ivf_nprobe) {
    size_t bytes = static_cast<size_t>(ncu_) * 8 * (2048 * 8 + 8 * 4) + 256;
    if (k > rows_) k = rows_;
    if (k >= 1 && (!fused_scan_serves(k) || (k > 1 && !fused_scan_serves(k - 1)))) {
        const size_t ksel = 2 * k - 1 < rows_ ? 2 * k - 1 : rows_;
        bytes += 4 * (rows_ + 2 * ksel) * 8 + 4 * (kSelectBins + 8) * 4 + select_sort_bytes(4, ksel);
    }
    Precision p;
    if (nq > kFusedBatchMax && exact_batch_precision(p)) {
        size_t most = 0;
        for (size_t kk = 1; kk <= k && kk <= 400 && kk <= rows_; ++kk) {
            const size_t b = plan_batch(kk, p, 0).workspace_bytes(nq, pdims_);
            most = b > most ? b : most;
        }
        bytes += most + nq;
    }
    if (ivf_nprobe) {
        const IvfIndex &ivf = build_ivf();
        if (ivf.nlist >= 2) {
            const size_t np = ivf_nprobe < ivf.nlist ? ivf_nprobe : ivf.nlist;
            bytes += nq * np * (8 + 4) + ivf_scan_scratch_bytes(nq, np, ivf.nlist, rows_, k);
            if (ivf_large_k(k)) {
                size_t g = 0, ld = 0;
                ivf_large_k_shape(ivf, nq, 1, k, g, ld);
                size_t g2 = 0, ld2 = 0;
                ivf_large_k_shape(ivf, nq, np, k, g2, ld2);
                bytes += g * ld2 * 8 + g * (2 * (2 * k - 1) * 8 + (kSelectBins + 16) * 4) +
                         select_sort_bytes(g, 2 * k - 1);
            }
            bytes += ivf.cindex->workspace_bytes(nq, np, 0);
        }
    }
    return bytes;
}

void GpuIndex::search_exact_any(Workspace &ws, const float *dq, size_t nq, size_t k,
                                uint64_t *d_idx, float *d_sim, hipStream_t s, bool heap_ties) {
    const uint32_t qbmax = scan_max_qb(static_cast<uint32_t>(k));
    Precision p;
    if (nq <= kFusedBatchMax || !exact_batch_precision(p) || k > 400 || qbmax == 0) {
        search_device(ws, dq, nq, k, d_idx, d_sim, s, heap_ties);
        return;
    }
    // MFMA scoring (BF16 where the row width allows, else FP32); the
    // certificate and the scan fallback make the results exact
    ws.reserve_xflags(nq);
    search_batch_mfma(ws, dq, nq, k, p, 0, d_idx, d_sim, ws.xflags, s);
    std::vector<uint8_t> h(nq);
    NNDEX_HIP_CHECK(hipMemcpyAsync(h.data(), ws.xflags, nq, hipMemcpyDeviceToHost, s));
    NNDEX_HIP_CHECK(hipStreamSynchronize(s));
    for (size_t q = 0; q < nq; ++q)
        if (h[q])
            search_device(ws, dq + q * pdims_, 1, k, d_idx + q * k, d_sim + q * k, s, heap_ties);
}

static void check_rows(size_t rows) {
    if (rows == 0 || rows >= 0xFFFFFFFFull)
        throw Error(NNDEX_ERR_INVALID_ARGUMENT,
                    "rows must be in [1, 2^32-1), got " + std::to_string(rows));
}

std::unique_ptr<GpuIndex> GpuIndex::from_storage(const float *storage, size_t rows,
                                                 size_t padded_dims, bool approx,
                                                 bool enable_cache, int64_t device_index) {
    check_rows(rows);
    if (padded_dims == 0 || padded_dims % 8 != 0)
        throw Error(NNDEX_ERR_INVALID_ARGUMENT, "padded_dims must be a positive multiple of 8");
    std::unique_ptr<GpuIndex> g(new GpuIndex());
    g->init_device(device_index);
    g->rows_ = rows;
    g->pdims_ = padded_dims;
    g->d4_ = padded_dims / 4;
    g->approx_ = approx;
    g->enable_cache_ = enable_cache;
    g->upload_tiled(storage, rows, padded_dims, false, false, padded_dims);
    return g;
}

std::unique_ptr<GpuIndex> GpuIndex::from_raw(const float *matrix, size_t rows, size_t dims,
                                             bool normalized, bool approx, bool enable_cache,
                                             int64_t device_index) {
    check_rows(rows);
    std::unique_ptr<GpuIndex> g(new GpuIndex());
    g->init_device(device_index);
    g->rows_ = rows;
    g->pdims_ = (dims + 7) / 8 * 8;
    g->d4_ = g->pdims_ / 4;
    g->approx_ = approx;
    g->enable_cache_ = enable_cache;
    g->upload_tiled(matrix, rows, dims, true, !normalized, dims);
    return g;
}

uint32_t GpuIndex::scan_tiles_per_wave() const {
    if (scan_tpw_) return scan_tpw_;
    // Default: about four tiles per wave, never more waves than 6 per SIMD.
    const uint32_t tiles = static_cast<uint32_t>(padded_rows(rows_) / kTileRows);
    uint32_t m = 4;
    while ((tiles + 4 * m - 1) / (4 * m) > ncu_ * 6) ++m;
    return m;
}

uint32_t GpuIndex::scan_grid() const {
    // Every wave owns the same number of 64-row tiles (the last wave may own
    // fewer), so no wave runs a tile longer than the others.
    const uint32_t tiles = static_cast<uint32_t>(padded_rows(rows_) / kTileRows);
    const uint32_t m = scan_tiles_per_wave();
    uint32_t g = (tiles + 4 * m - 1) / (4 * m);
    // at most 8 blocks per CU and 2048 blocks (the final merge's LDS table)
    const uint32_t cap = ncu_ * 8 < 2048 ? ncu_ * 8 : 2048;
    if (g > cap) g = cap;
    return g ? g : 1;
}

std::string GpuIndex::autotune(int reps) {
    NNDEX_HIP_CHECK(hipSetDevice(device_));
    std::lock_guard<std::mutex> lock(mu_);
    const uint32_t tiles = static_cast<uint32_t>(padded_rows(rows_) / kTileRows);
    std::vector<float> q(pdims_, 0.f);
    for (size_t i = 0; i < pdims_; ++i) q[i] = ((i * 2654435761u) % 1000) / 1000.0f - 0.5f;
    DeviceBuffer dq, di, ds;
    dq.reserve(pdims_ * sizeof(float));
    di.reserve(16 * sizeof(uint64_t));
    ds.reserve(16 * sizeof(float));
    NNDEX_HIP_CHECK(hipMemcpy(dq.ptr, q.data(), pdims_ * sizeof(float), hipMemcpyHostToDevice));
    hipEvent_t e0, e1;
    NNDEX_HIP_CHECK(hipEventCreate(&e0));
    NNDEX_HIP_CHECK(hipEventCreate(&e1));
    std::string report;
    float best_ms = 1e30f;
    uint32_t best_m = 0, best_f = kScanNonTemporal;
    std::vector<uint32_t> flag_set;
    for (uint32_t grp : {0u, 1u, 2u})
        for (uint32_t f : {kScanNonTemporal, 0u, kScanNonTemporal | kScanContiguous,
                           kScanContiguous})
            flag_set.push_back(f | (grp << kScanGroupShift));
    for (uint32_t f : flag_set) {
        for (uint32_t m : {2u, 3u, 4u, 6u, 8u, 12u, 14u, 16u, 18u, 20u, 24u, 32u, 48u}) {
            if ((tiles + 4 * m - 1) / (4 * m) > ncu_ * 8) continue;
            scan_tpw_ = m;
            scan_flags_ = f;
            auto ws = make_workspace(1, 10);
            for (int i = 0; i < 3; ++i)
                search_device(*ws, static_cast<const float *>(dq.ptr), 1, 10,
                              static_cast<uint64_t *>(di.ptr), static_cast<float *>(ds.ptr),
                              stream_);
            NNDEX_HIP_CHECK(hipEventRecord(e0, stream_));
            for (int i = 0; i < reps; ++i)
                search_device(*ws, static_cast<const float *>(dq.ptr), 1, 10,
                              static_cast<uint64_t *>(di.ptr), static_cast<float *>(ds.ptr),
                              stream_);
            NNDEX_HIP_CHECK(hipEventRecord(e1, stream_));
            NNDEX_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0;
            NNDEX_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            ms /= reps;
            report += "scan flags=" + std::to_string(f) + " tiles_per_wave=" + std::to_string(m) +
                      " grid=" + std::to_string(scan_grid()) +
                      " us=" + std::to_string(ms * 1e3f) + "\n";
            if (ms < best_ms) {
                best_ms = ms;
                best_m = m;
                best_f = f;
            }
        }
    }
    scan_flags_ = best_f;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    scan_tpw_ = best_m;
    ws_.reset();  // grids changed; the host path re-creates its workspace
    report += "chosen flags=" + std::to_string(best_f) + " tiles_per_wave=" +
              std::to_string(best_m) + "\n";
    return report;
}

std::string GpuIndex::autotune_batch(int reps) {
    NNDEX_HIP_CHECK(hipSetDevice(device_));
    std::lock_guard<std::mutex> lock(mu_);
    ensure_rowmajor(stream_);
    // the index's first rows as queries (row-major FP32, already normalised)
    const size_t nq = rows_ < 2048 ? rows_ : 2048, k = rows_ < 10 ? rows_ : 10;
    const float *dq = reinterpret_cast<const float *>(xrow_);
    DeviceBuffer di, ds, df;
    di.reserve(nq * k * sizeof(uint64_t));
    ds.reserve(nq * k * sizeof(float));
    df.reserve(nq);
    auto ws = make_workspace(nq, k);
    hipEvent_t e0, e1;
    NNDEX_HIP_CHECK(hipEventCreate(&e0));
    NNDEX_HIP_CHECK(hipEventCreate(&e1));
    std::string report;
    for (Precision p : {Precision::F32, Precision::BF16, Precision::FP8}) {
        if (!mfma_supported(p)) continue;
        const int elem = static_cast<int>(p);
        const uint32_t nv = mfma_variants(elem, static_cast<uint32_t>(mfma_units(pdims_, p)));
        float best_ms = 1e30f;
        int best_v = 0;
        for (uint32_t v = 0; v < nv; ++v) {
            mfma_variant_[elem] = static_cast<int>(v);
            auto run = [&] {
                search_batch_mfma(*ws, dq, nq, k, p, 0, static_cast<uint64_t *>(di.ptr),
                                  static_cast<float *>(ds.ptr), static_cast<uint8_t *>(df.ptr),
                                  stream_);
            };
            run();
            NNDEX_HIP_CHECK(hipEventRecord(e0, stream_));
            for (int i = 0; i < reps; ++i) run();
            NNDEX_HIP_CHECK(hipEventRecord(e1, stream_));
            NNDEX_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0;
            NNDEX_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            ms /= reps;
            static const char *names[] = {"f32", "bf16", "fp8"};
            report += std::string("batch ") + names[elem] + " variant=" + std::to_string(v) +
                      " rows_per_block=" +
                      std::to_string(mfma_block_rows(elem,
                                                     static_cast<uint32_t>(mfma_units(pdims_, p)),
                                                     false, static_cast<int>(v))) +
                      " ms=" + std::to_string(ms) + "\n";
            if (ms < best_ms) {
                best_ms = ms;
                best_v = static_cast<int>(v);
            }
        }
        mfma_variant_[elem] = best_v;
        report += std::string("chosen ") + std::to_string(elem) + " variant=" +
                  std::to_string(best_v) + "\n";
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return report;
}

std::unique_ptr<Workspace> GpuIndex::make_workspace(size_t max_queries, size_t max_k) const {
    NNDEX_HIP_CHECK(hipSetDevice(device_));
    std::unique_ptr<Workspace> ws(new Workspace());
    ws->device = device_;
    ws->max_queries = max_queries;
    ws->max_k = max_k;
    ws->grid = scan_grid();
    // QB * k < 2048 for every fused-scan variant; room for the largest grid.
    const size_t gmax = static_cast<size_t>(ncu_) * 8;
    ws->part_elems = 2048 * gmax;
    NNDEX_HIP_CHECK(hipMalloc(&ws->part, ws->part_elems * sizeof(uint64_t)));
    NNDEX_HIP_CHECK(hipMalloc(&ws->part_count, 8 * gmax * sizeof(uint32_t)));
    NNDEX_HIP_CHECK(hipMalloc(&ws->theta, 8 * sizeof(unsigned long long)));
    NNDEX_HIP_CHECK(hipMalloc(&ws->ticket, sizeof(uint32_t)));
    NNDEX_HIP_CHECK(hipMemset(ws->part_count, 0, 8 * gmax * sizeof(uint32_t)));
    NNDEX_HIP_CHECK(hipMemset(ws->theta, 0, 8 * sizeof(unsigned long long)));
    NNDEX_HIP_CHECK(hipMemset(ws->ticket, 0, sizeof(uint32_t)));
    return ws;
}

// Keys the fused scan selects per query for k results: the reference heap's
// choice among rows tied at the k-th score is recovered from the canonical
// top 2k-1 (k above ~990: canonical order).
static size_t scan_select_k(size_t k, size_t rows, bool heap_ties) {
    size_t ksel = k;
    if (heap_ties && k > 1 && k < rows) {
        const size_t k2 = 2 * k - 1 < rows ? 2 * k - 1 : rows;
        if (scan_max_qb(static_cast<uint32_t>(k2)) > 0) ksel = k2;
    }
    return ksel;
}

size_t GpuIndex::scan_part_keys(size_t k, bool heap_ties) const {
    return static_cast<size_t>(scan_grid()) * scan_select_k(k, rows_, heap_ties);
}

bool GpuIndex::fused_scan_serves(size_t k, bool heap_ties) const {
    if (heap_ties && k > 1 && k < rows_) {
        const size_t k2 = 2 * k - 1 < rows_ ? 2 * k - 1 : rows_;
        return scan_max_qb(static_cast<uint32_t>(k2)) > 0;
    }
    return scan_max_qb(static_cast<uint32_t>(k)) > 0;
}

void GpuIndex::search_large_k(Workspace &ws, const float *dq, size_t nq, size_t k,
                              uint64_t *d_idx, float *d_sim, hipStream_t s, bool heap_ties) const {
    if (k == 0 || k > rows_) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "k must be in [1, rows]");
    // the canonical top 2k - 1 holds every row the heap's tie rule can need
    const size_t ksel = heap_ties && k > 1 && k < rows_ ? (2 * k - 1 < rows_ ? 2 * k - 1 : rows_)
                                                        : k;
    constexpr size_t group = 4;
    ws.reserve_select(group, rows_, ksel);
    for (size_t q0 = 0; q0 < nq; q0 += group) {
        const uint32_t n = static_cast<uint32_t>(nq - q0 < group ? nq - q0 : group);
        launch_score_keys(x_, static_cast<uint32_t>(rows_), static_cast<uint32_t>(d4_),
                          dq + q0 * pdims_, n, ws.sel_keys, ws.sel_ld, ncu_, s);
        const SelectBuffers b = ws.select_buffers(n, static_cast<uint32_t>(ksel), nullptr,
                                                  static_cast<uint32_t>(rows_));
        select_top(b, s);
        launch_finish_exact(b, static_cast<uint32_t>(k), heap_ties, d_idx + q0 * k,
                            d_sim + q0 * k, k, s);
    }
}

void GpuIndex::search_device(Workspace &ws, const float *dq, size_t nq, size_t k, uint64_t *d_idx,
                             float *d_sim, hipStream_t s, bool heap_ties) const {
    if (k > ws.max_k && k <= rows_)
        throw Error(NNDEX_ERR_INVALID_ARGUMENT, "k exceeds the workspace's max_k");
    if (static_cast<size_t>(scan_grid()) * 2048 > ws.part_elems)
        throw Error(NNDEX_ERR_INVALID_ARGUMENT,
                    "workspace was created for a smaller launch shape; create a new one");
    if (k >= 1 && k <= rows_ && !fused_scan_serves(k, heap_ties)) {
        search_large_k(ws, dq, nq, k, d_idx, d_sim, s, heap_ties);
        return;
    }
    ScanScratch sc;
    sc.part = ws.part;
    sc.part_elems = ws.part_elems;
    sc.part_count = ws.part_count;
    sc.theta = ws.theta;
    sc.ticket = ws.ticket;
    search_device(sc, dq, nq, k, d_idx, d_sim, s, heap_ties);
}

void GpuIndex::search_device(const ScanScratch &sc, const float *dq, size_t nq, size_t k,
                             uint64_t *d_idx, float *d_sim, hipStream_t s, bool heap_ties) const {
    if (k == 0 || k > rows_)
        throw Error(NNDEX_ERR_INVALID_ARGUMENT, "k must be in [1, rows]");
    const size_t ksel = scan_select_k(k, rows_, heap_ties);
    if (!fused_scan_serves(k, heap_ties))
        throw Error(NNDEX_ERR_UNSUPPORTED,
                    "k=" + std::to_string(k) +
                        " exceeds the fused scan (k <= 992, or k = rows <= 1984); larger k "
                        "needs a workspace");
    uint32_t qbmax = scan_max_qb(static_cast<uint32_t>(ksel));
    if (qbmax == 0)
        throw Error(NNDEX_ERR_UNSUPPORTED, "k=" + std::to_string(k) + " exceeds the fused scan");
    // every launch publishes at most grid x ksel keys per query; take as many
    // queries per launch as the scratch holds (results do not depend on it)
    const size_t per_query = static_cast<size_t>(scan_grid()) * ksel;
    while (qbmax > 1 && qbmax * per_query > sc.part_elems) qbmax >>= 1;
    if (per_query > sc.part_elems)
        throw Error(NNDEX_ERR_UNSUPPORTED,
                    "k=" + std::to_string(k) + " needs " + std::to_string(per_query) +
                        " block-list keys per launch, above the scratch's " +
                        std::to_string(sc.part_elems));
    for (size_t q0 = 0; q0 < nq; q0 += qbmax) {
        uint32_t nvalid = static_cast<uint32_t>(nq - q0 < qbmax ? nq - q0 : qbmax);
        uint32_t qb = 1;
        while (qb < nvalid) qb <<= 1;
        ScanArgs a{};
        a.x = x_;
        a.rows = static_cast<uint32_t>(rows_);
        a.d4 = static_cast<uint32_t>(d4_);
        a.q = reinterpret_cast<const float4 *>(dq + q0 * pdims_);
        a.q_stride4 = static_cast<uint32_t>(d4_);
        a.nvalid = nvalid;
        a.k = static_cast<uint32_t>(ksel);
        a.k_out = static_cast<uint32_t>(k);
        a.part = sc.part;
        a.part_count = sc.part_count;
        a.theta = sc.theta;
        a.ticket = sc.ticket;
        a.out_idx = d_idx + q0 * k;
        a.out_sim = d_sim + q0 * k;
        a.out_stride = static_cast<uint32_t>(k);
        a.flags = scan_flags_;
        const uint32_t grid = scan_grid();
        if (static_cast<size_t>(qb) * grid * ksel > sc.part_elems)
            throw Error(NNDEX_ERR_INVALID_ARGUMENT, "fused scan block lists exceed the scratch");
        launch_exact_scan(a, qb, grid, s);
    }
}

uint64_t GpuIndex::query_hash(const float *values, size_t n, size_t k, size_t query_rows,
                              bool approx) const {
    // FNV-1a over the fields src/gpu.rs:787-797 hashes.
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void *p, size_t bytes) {
        const unsigned char *c = static_cast<const unsigned char *>(p);
        for (size_t i = 0; i < bytes; ++i) {
            h ^= c[i];
            h *= 1099511628211ull;
        }
    };
    uint64_t f[5] = {rows_, pdims_, k, query_rows, approx ? 1ull : 0ull};
    mix(f, sizeof(f));
    mix(values, n * sizeof(float));
    return h;
}

std::vector<std::vector<Neighbor>> GpuIndex::run_host(const float *queries, size_t query_rows,
                                                      size_t k, size_t raw_dims, int mode,
                                                      bool approx) {
    NNDEX_HIP_CHECK(hipSetDevice(device_));
    // dispatch_*_or_approx (src/gpu.rs:359-367, :404-432)
    const IvfIndex *ivf = approx ? approx_index() : nullptr;
    if (!ws_ || ws_->max_k < k) ws_ = make_workspace(8, k < 10 ? 10 : k);
    const size_t qbytes = query_rows * pdims_ * sizeof(float);
    const size_t in_bytes = query_rows * (raw_dims ? raw_dims : pdims_) * sizeof(float);
    const size_t obytes = query_rows * k;
    d_q_.reserve(qbytes);
    d_idx_.reserve(obytes * sizeof(uint64_t));
    d_sim_.reserve(obytes * sizeof(float));
    h_q_.reserve(in_bytes > sizeof(uint32_t) ? in_bytes : sizeof(uint32_t));
    h_idx_.reserve(obytes * sizeof(uint64_t));
    h_sim_.reserve(obytes * sizeof(float));
    std::memcpy(h_q_.ptr, queries, in_bytes);
    // raw queries are checked and prepared on the device; their flags come
    // back with the results (one synchronisation for the whole call), and a
    // failed row was zeroed, so the search over it is well defined
    uint32_t *h_flags = nullptr;
    if (raw_dims) {
        d_raw_.reserve(in_bytes);
        d_flags_.reserve(sizeof(uint32_t));
        h_flag_.reserve(sizeof(uint32_t));
        h_flags = static_cast<uint32_t *>(h_flag_.ptr);
        launch_zero_u32(static_cast<uint32_t *>(d_flags_.ptr), 1, stream_);
        NNDEX_HIP_CHECK(
            hipMemcpyAsync(d_raw_.ptr, h_q_.ptr, in_bytes, hipMemcpyHostToDevice, stream_));
        launch_normalize_rowmajor(static_cast<const float *>(d_raw_.ptr),
                                  static_cast<uint32_t>(query_rows), static_cast<uint32_t>(raw_dims),
                                  static_cast<uint32_t>(pdims_), mode,
                                  static_cast<float *>(d_q_.ptr),
                                  static_cast<uint32_t *>(d_flags_.ptr), nullptr, stream_);
        NNDEX_HIP_CHECK(hipMemcpyAsync(h_flags, d_flags_.ptr, sizeof(uint32_t),
                                       hipMemcpyDeviceToHost, stream_));
    } else {
        NNDEX_HIP_CHECK(
            hipMemcpyAsync(d_q_.ptr, h_q_.ptr, qbytes, hipMemcpyHostToDevice, stream_));
    }
    std::vector<uint64_t> lens;
    if (ivf) {
        d_len_.reserve(query_rows * sizeof(uint64_t));
        search_ivf_device(*ws_, static_cast<const float *>(d_q_.ptr), query_rows, k, 0,
                          static_cast<uint64_t *>(d_idx_.ptr), static_cast<float *>(d_sim_.ptr),
                          static_cast<uint64_t *>(d_len_.ptr), stream_);
        lens.resize(query_rows);
        NNDEX_HIP_CHECK(hipMemcpyAsync(lens.data(), d_len_.ptr, query_rows * sizeof(uint64_t),
                                       hipMemcpyDeviceToHost, stream_));
    } else {
        search_exact_any(*ws_, static_cast<const float *>(d_q_.ptr), query_rows, k,
                         static_cast<uint64_t *>(d_idx_.ptr), static_cast<float *>(d_sim_.ptr),
                         stream_);
    }
    NNDEX_HIP_CHECK(hipMemcpyAsync(h_idx_.ptr, d_idx_.ptr, obytes * sizeof(uint64_t),
                                   hipMemcpyDeviceToHost, stream_));
    NNDEX_HIP_CHECK(hipMemcpyAsync(h_sim_.ptr, d_sim_.ptr, obytes * sizeof(float),
                                   hipMemcpyDeviceToHost, stream_));
    if (hipStreamSynchronize(stream_) != hipSuccess)
        throw Error(NNDEX_ERR_GPU_MAP_READ_FAILED, "failed to map GPU buffer for reading");
    if (h_flags) throw_flags(*h_flags);
    std::vector<std::vector<Neighbor>> out(query_rows);
    const uint64_t *hi = static_cast<const uint64_t *>(h_idx_.ptr);
    const float *hs = static_cast<const float *>(h_sim_.ptr);
    for (size_t q = 0; q < query_rows; ++q) {
        const size_t n = ivf ? static_cast<size_t>(lens[q]) : k;
        out[q].resize(n);
        for (size_t i = 0; i < n; ++i) out[q][i] = {hi[q * k + i], hs[q * k + i]};
    }
    return out;
}

std::vector<Neighbor> GpuIndex::search(const float *query, size_t k, bool approx) {
    std::lock_guard<std::mutex> lock(mu_);
    const size_t capped = k < rows_ ? k : rows_;
    if (enable_cache_) {
        uint64_t key = query_hash(query, pdims_, k, 1, approx);
        auto it = cache_single_.find(key);
        if (it != cache_single_.end()) return it->second;
        auto res = std::move(run_host(query, 1, capped, 0, kQueryPad, approx)[0]);
        if (cache_single_.size() >= kSingleCacheCapacity) cache_single_.clear();
        cache_single_.emplace(key, res);
        return res;
    }
    return std::move(run_host(query, 1, capped, 0, kQueryPad, approx)[0]);
}

std::vector<std::vector<Neighbor>> GpuIndex::search_batch(const float *queries, size_t query_rows,
                                                          size_t k, bool approx) {
    std::lock_guard<std::mutex> lock(mu_);
    const size_t capped = k < rows_ ? k : rows_;
    if (query_rows == 0) return {};
    if (enable_cache_) {
        uint64_t key = query_hash(queries, query_rows * pdims_, k, query_rows, approx);
        auto it = cache_batch_.find(key);
        if (it != cache_batch_.end()) return it->second;
        auto res = run_host(queries, query_rows, capped, 0, kQueryPad, approx);
        if (cache_batch_.size() >= kBatchCacheCapacity) cache_batch_.clear();
        cache_batch_.emplace(key, res);
        return res;
    }
    return run_host(queries, query_rows, capped, 0, kQueryPad, approx);
}

std::vector<std::vector<Neighbor>> GpuIndex::search_batch_raw(const float *queries,
                                                              size_t query_rows, size_t dims,
                                                              bool normalize, size_t k,
                                                              bool approx) {
    std::lock_guard<std::mutex> lock(mu_);
    const size_t capped = k < rows_ ? k : rows_;
    if (query_rows == 0) return {};
    if (enable_cache_) {
        // Normalisation is deterministic, so keying on the raw values (plus
        // the mode) hits exactly when a key over the normalised values would.
        uint64_t key = query_hash(queries, query_rows * dims, k, query_rows, approx) ^
                       (normalize ? 0x9e3779b97f4a7c15ull : 0ull);
        auto it = cache_batch_.find(key);
        if (it != cache_batch_.end()) return it->second;
        auto res = run_host(queries, query_rows, capped, dims,
                            normalize ? kQueryBatch : kQueryPad, approx);
        if (cache_batch_.size() >= kBatchCacheCapacity) cache_batch_.clear();
        cache_batch_.emplace(key, res);
        return res;
    }
    return run_host(queries, query_rows, capped, dims, normalize ? kQueryBatch : kQueryPad,
                    approx);
}

std::vector<Neighbor> GpuIndex::search_raw(const float *query, size_t dims, bool normalize,
                                           size_t k, bool approx) {
    std::lock_guard<std::mutex> lock(mu_);
    const size_t capped = k < rows_ ? k : rows_;
    const int mode = normalize ? kQuerySingle : kQueryPad;
    if (enable_cache_) {
        // prepare_query is deterministic: a key over the raw values and the
        // mode hits exactly when one over the prepared query would
        uint64_t key = query_hash(query, dims, k, 1, approx) ^
                       (normalize ? 0x5bd1e9955bd1e995ull : 0x27d4eb2f165667c5ull);
        auto it = cache_single_.find(key);
        if (it != cache_single_.end()) return it->second;
        auto res = std::move(run_host(query, 1, capped, dims, mode, approx)[0]);
        if (cache_single_.size() >= kSingleCacheCapacity) cache_single_.clear();
        cache_single_.emplace(key, res);
        return res;
    }
    return std::move(run_host(query, 1, capped, dims, mode, approx)[0]);
}

// ---------------------------------------------------------------------------
// IVF prefilter
// ---------------------------------------------------------------------------

std::unique_ptr<GpuIndex> GpuIndex::from_device_rows(const float *d_rows, size_t rows,
                                                     size_t padded_dims, int device) {
    check_rows(rows);
    std::unique_ptr<GpuIndex> g(new GpuIndex());
    g->init_device(device);
    g->rows_ = rows;
    g->pdims_ = padded_dims;
    g->d4_ = padded_dims / 4;
    g->enable_cache_ = false;
    const size_t tiles = padded_rows(rows) / kTileRows;
    NNDEX_HIP_CHECK(hipMalloc(&g->x_, tiles * g->d4_ * kTileRows * sizeof(float4)));
    DeviceBuffer flags;
    flags.reserve(sizeof(uint32_t));
    launch_zero_u32(static_cast<uint32_t *>(flags.ptr), 1, g->stream_);
    launch_to_tiled(d_rows, static_cast<uint32_t>(rows), static_cast<uint32_t>(padded_dims),
                    static_cast<uint32_t>(g->d4_), g->x_, static_cast<uint32_t *>(flags.ptr),
                    g->stream_);
    NNDEX_HIP_CHECK(hipMalloc(&g->xmax_sq_, sizeof(unsigned long long)));
    NNDEX_HIP_CHECK(hipMemsetAsync(g->xmax_sq_, 0, sizeof(unsigned long long), g->stream_));
    launch_max_norm(g->x_, static_cast<uint32_t>(rows), static_cast<uint32_t>(g->d4_),
                    g->xmax_sq_, g->stream_);
    NNDEX_HIP_CHECK(hipStreamSynchronize(g->stream_));
    return g;
}

void GpuIndex::set_ivf_params(size_t nlist, size_t nprobe, size_t iters) {
    std::lock_guard<std::mutex> l(ivf_mu_);
    if (nlist > rows_ || nlist == 1 || nlist > 16384)
        throw Error(NNDEX_ERR_INVALID_ARGUMENT, "nlist must be 0 or in [2, min(rows, 16384)]");
    ivf_nlist_ = nlist;
    ivf_nprobe_ = nprobe;
    ivf_iters_ = iters;
}

const IvfIndex &GpuIndex::build_ivf() {
    std::lock_guard<std::mutex> l(ivf_mu_);
    if (ivf_) return *ivf_;
    NNDEX_HIP_CHECK(hipSetDevice(device_));
    std::unique_ptr<IvfIndex> ivf(new IvfIndex());
    const size_t nlist = ivf_nlist_ ? ivf_nlist_ : ivf_choose_nlist(rows_);
    const size_t iters = ivf_iters_ ? ivf_iters_ : kKmeansMaxIters;
    if (nlist < 2) {
        // ApproxIndex::build's single-list form (approx.rs:50-60): never used
        // for searching (provides_speedup is false), kept for the export.
        ivf->rows = static_cast<uint32_t>(rows_);
        ivf->pdims = static_cast<uint32_t>(pdims_);
        ivf->nlist = 1;
        ivf_ = std::move(ivf);
        return *ivf_;
    }
    unsigned long long bits = 0, ebits = 0;
    NNDEX_HIP_CHECK(hipMemcpy(&bits, xmax_sq_, sizeof(bits), hipMemcpyDeviceToHost));
    const double x_max = std::sqrt(__builtin_bit_cast(double, bits));
    double x_err = 0.0;
    {  // BF16 copy: the k-means++ pre-test, and the assignment above 512 dims
        ensure_bf16(stream_);
        NNDEX_HIP_CHECK(hipStreamSynchronize(stream_));
        NNDEX_HIP_CHECK(hipMemcpy(&ebits, xerr_sq_ + 1, sizeof(ebits), hipMemcpyDeviceToHost));
        x_err = std::sqrt(__builtin_bit_cast(double, ebits));
    }
    Fp8Rows fp8;
    {  // FP8 copy: the k-means++ pre-test (half the bytes of the BF16 one)
        ensure_fp8(stream_);
        unsigned long long e8 = 0;
        NNDEX_HIP_CHECK(hipMemcpy(&e8, xerr8_sq_, sizeof(e8), hipMemcpyDeviceToHost));
        fp8.codes = xfp8_;
        fp8.scale = xscale_;
        fp8.err = std::sqrt(__builtin_bit_cast(double, e8));
    }
    ivf_build(*ivf, x_, xbf16_, static_cast<uint32_t>(rows_), static_cast<uint32_t>(pdims_),
              static_cast<uint32_t>(nlist), static_cast<uint32_t>(iters), x_max, x_err, fp8, ncu_,
              stream_);
    ivf->cindex = from_device_rows(ivf->centroids, nlist, pdims_, device_);
    ivf_ = std::move(ivf);
    return *ivf_;
}

const IvfIndex &GpuIndex::load_ivf(size_t nlist, const float *centroids, const uint64_t *offsets,
                                   const uint64_t *orig, const uint64_t *assign) {
    std::lock_guard<std::mutex> l(ivf_mu_);
    if (ivf_) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "an IVF index already exists on this index");
    if (!centroids || !offsets || !orig)
        throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
    if (nlist < 2 || nlist > rows_ || nlist > 16384)
        throw Error(NNDEX_ERR_INVALID_ARGUMENT, "nlist must be in [2, min(rows, 16384)]");
    if (pdims_ > 16384)
        throw Error(NNDEX_ERR_UNSUPPORTED, "the GPU IVF supports padded dims <= 16384");
    for (size_t i = 0; i < nlist * pdims_; ++i)
        if (!std::isfinite(centroids[i]))
            throw Error(NNDEX_ERR_INVALID_ARGUMENT, "IVF centroids must be finite");
    if (offsets[0] != 0 || offsets[nlist] != rows_)
        throw Error(NNDEX_ERR_INVALID_ARGUMENT, "IVF offsets must run from 0 to rows");
    std::vector<uint32_t> off(nlist + 1), ord(rows_), asg(rows_);
    for (size_t c = 0; c <= nlist; ++c) {
        if (c && offsets[c] < offsets[c - 1])
            throw Error(NNDEX_ERR_INVALID_ARGUMENT, "IVF offsets must be non-decreasing");
        off[c] = static_cast<uint32_t>(offsets[c]);
    }
    std::vector<uint8_t> seen(rows_, 0);
    for (size_t c = 0; c < nlist; ++c)
        for (size_t p = off[c]; p < off[c + 1]; ++p) {
            const uint64_t r = orig[p];
            if (r >= rows_ || seen[r])
                throw Error(NNDEX_ERR_INVALID_ARGUMENT,
                            "IVF original_indices must be a permutation of the rows");
            if (assign && assign[r] != c)
                throw Error(NNDEX_ERR_INVALID_ARGUMENT,
                            "IVF assignments disagree with the offsets and original_indices");
            seen[r] = 1;
            ord[p] = static_cast<uint32_t>(r);
            asg[r] = static_cast<uint32_t>(c);
        }
    NNDEX_HIP_CHECK(hipSetDevice(device_));
    std::unique_ptr<IvfIndex> ivf(new IvfIndex());
    ivf_load(*ivf, x_, static_cast<uint32_t>(rows_), static_cast<uint32_t>(pdims_),
             static_cast<uint32_t>(nlist), centroids, off.data(), ord.data(), asg.data(), stream_);
    ivf->cindex = from_device_rows(ivf->centroids, nlist, pdims_, device_);
    ivf_ = std::move(ivf);
    return *ivf_;
}

const IvfIndex *GpuIndex::approx_index() {
    if (!approx_) return nullptr;
    const IvfIndex &ivf = build_ivf();
    return ivf.provides_speedup() ? &ivf : nullptr;
}

void GpuIndex::search_ivf_device(Workspace &ws, const float *dq, size_t nq, size_t k,
                                 size_t nprobe, uint64_t *d_idx, float *d_sim, uint64_t *d_len,
                                 hipStream_t s) {
    if (k == 0 || k > rows_) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "k must be in [1, rows]");
    const IvfIndex &ivf = build_ivf();
    if (ivf.nlist < 2)
        throw Error(NNDEX_ERR_UNSUPPORTED, "the IVF index has a single list (rows < 256)");
    if (nq == 0) return;
    if (nprobe == 0) nprobe = ivf_nprobe_;
    if (nprobe == 0) nprobe = ivf_choose_nprobe(ivf.nlist, k, rows_);
    if (nprobe > ivf.nlist) nprobe = ivf.nlist;
    if (!ws.sub || ws.sub->max_k < nprobe)
        ws.sub = ivf.cindex->make_workspace(8, nprobe < 10 ? 10 : nprobe);
    if (ivf_large_k(k)) {
        // every probed row's key, then the top 2k - 1 per query by radix
        // select, in groups of queries whose key arrays fit 256 MiB
        size_t group = 0, ld = 0;
        ivf_large_k_shape(ivf, nq, nprobe, k, group, ld);
        const size_t ksel = 2 * k - 1;
        ws.reserve_select(group, ld, ksel);
        for (size_t q0 = 0; q0 < nq; q0 += group) {
            const size_t n = nq - q0 < group ? nq - q0 : group;
            uint32_t qb, nchunk;
            ivf_scan_shape(n, nprobe, ivf.nlist, rows_, 1, qb, nchunk);
            const size_t pairs = n * nprobe;
            ws.reserve_ivf(pairs, ivf.nlist, 0, ivf_sort_bytes(pairs, ivf.nlist));
            ivf.cindex->search_exact_any(*ws.sub, dq + q0 * pdims_, n, nprobe, ws.ivf_probe,
                                         ws.ivf_probe_sim, s, false);
            IvfScanArgs a = ivf_scan_args(ivf, ws, dq + q0 * pdims_, n, k, nprobe, 1, qb, nchunk,
                                          d_idx + q0 * k, d_sim + q0 * k, d_len + q0);
            a.keys_out = ws.sel_keys;
            a.keys_ld = ws.sel_ld;
            a.totals = ws.sel_n;
            ivf_scan(a, s);
            const SelectBuffers b =
                ws.select_buffers(static_cast<uint32_t>(n), static_cast<uint32_t>(ksel), ws.sel_n, 0);
            select_top(b, s);
            ivf_finish_large_k(a, b, static_cast<uint32_t>(k), s);
        }
        return;
    }
    uint32_t qb, nchunk;
    const uint32_t kk = ivf_select_k(k);
    ivf_scan_shape(nq, nprobe, ivf.nlist, rows_, kk, qb, nchunk);
    const size_t pairs = nq * nprobe;
    ws.reserve_ivf(pairs, ivf.nlist, pairs * nchunk * kk, ivf_sort_bytes(pairs, ivf.nlist));
    // candidate_indices (approx.rs:110-138): top-nprobe centroids by score
    // the probe ranks centroids in (similarity desc, index asc) order
    ivf.cindex->search_exact_any(*ws.sub, dq, nq, nprobe, ws.ivf_probe, ws.ivf_probe_sim, s,
                                 false);
    ivf_scan(ivf_scan_args(ivf, ws, dq, nq, k, nprobe, kk, qb, nchunk, d_idx, d_sim, d_len), s);
}

void GpuIndex::ivf_large_k_shape(const IvfIndex &ivf, size_t nq, size_t nprobe, size_t k,
                                 size_t &group, size_t &ld) const {
    ld = ivf.max_probe_rows(nprobe);
    if (ld < 1) ld = 1;
    group = (size_t(256) << 20) / (ld * sizeof(uint64_t));
    const size_t by_k = (size_t(256) << 20) / ((2 * k - 1) * 2 * sizeof(uint64_t));
    if (group > by_k) group = by_k;
    if (group < 1) group = 1;
    if (group > 64) group = 64;
    if (group > nq) group = nq;
}

IvfScanArgs GpuIndex::ivf_scan_args(const IvfIndex &ivf, Workspace &ws, const float *dq,
                                    size_t nq, size_t k, size_t nprobe, uint32_t kk, uint32_t qb,
                                    uint32_t nchunk, uint64_t *d_idx, float *d_sim,
                                    uint64_t *d_len) const {
    const size_t pairs = nq * nprobe;
    IvfScanArgs a{};
    a.x = ivf.xr;
    a.d4 = static_cast<uint32_t>(d4_);
    a.q = dq;
    a.probe = ws.ivf_probe;
    a.nq = static_cast<uint32_t>(nq);
    a.nprobe = static_cast<uint32_t>(nprobe);
    a.nlist = ivf.nlist;
    a.k = static_cast<uint32_t>(k);
    a.kk = kk;
    a.qb = qb;
    a.nchunk = nchunk;
    a.offsets = ivf.offsets;
    a.orig = ivf.orig;
    a.pair_key = ws.ivf_pair_u32;
    a.pair_val = ws.ivf_pair_u32 + pairs;
    a.sorted_key = ws.ivf_pair_u32 + 2 * pairs;
    a.sorted_pair = ws.ivf_pair_u32 + 3 * pairs;
    a.pair_base = ws.ivf_pair_u32 + 4 * pairs;
    const size_t L1 = ivf.nlist + 1;
    a.pair_cnt = ws.ivf_list_u32;
    a.pair_off = ws.ivf_list_u32 + L1;
    a.grp_cnt = ws.ivf_list_u32 + 2 * L1;
    a.grp_off = ws.ivf_list_u32 + 3 * L1;
    a.part = ws.ivf_part;
    a.sort_tmp = ws.ivf_sort_tmp;
    a.sort_bytes = ws.ivf_sort_bytes;
    a.out_idx = d_idx;
    a.out_sim = d_sim;
    a.out_len = d_len;
    return a;
}

}  // namespace nndex

// ===========================================================================
// C ABI: backend boundary, device-resident search, helpers
// ===========================================================================
using namespace nndex;

struct nndex_workspace {
    std::unique_ptr<Workspace> impl;
    const GpuIndex *owner;
};

extern "C" {

const char *nndex_last_error(void) { return g_last_error.c_str(); }
const char *nndex_version(void) { return "nndex-mi355x 0.1.0 (gfx950)"; }

nndex_status nndex_gpu_index_new(const float *storage, size_t rows, size_t padded_dims,
                                 int32_t approx, int32_t enable_cache, int64_t device_index,
                                 nndex_gpu_index **out) {
    return guard([&] {
        if (!storage || !out) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        auto h = new nndex_gpu_index();
        try {
            h->impl = GpuIndex::from_storage(storage, rows, padded_dims, approx != 0,
                                             enable_cache != 0, device_index);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

void nndex_gpu_index_free(nndex_gpu_index *index) { delete index; }
// Handles returned by nndex_backend_handle belong to their nndex_index and
// must not be passed to nndex_gpu_index_free.

size_t nndex_gpu_index_rows(const nndex_gpu_index *index) {
    return index ? index->impl->rows() : 0;
}
size_t nndex_gpu_index_padded_dims(const nndex_gpu_index *index) {
    return index ? index->impl->padded_dims() : 0;
}

nndex_status nndex_gpu_index_search(nndex_gpu_index *index, const float *query, size_t k,
                                    int32_t approx, uint64_t *out_indices,
                                    float *out_similarities, size_t *out_len) {
    return guard([&] {
        if (!index || !query || !out_indices || !out_similarities || !out_len)
            throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        if (k == 0) throw Error(NNDEX_ERR_INVALID_TOP_K, "top-k must be greater than zero");
        auto res = index->impl->search(query, k, approx != 0);
        for (size_t i = 0; i < res.size(); ++i) {
            out_indices[i] = res[i].index;
            out_similarities[i] = res[i].similarity;
        }
        *out_len = res.size();
    });
}

nndex_status nndex_gpu_index_search_batch(nndex_gpu_index *index, const float *queries,
                                          size_t query_rows, size_t k, int32_t approx,
                                          uint64_t *out_indices, float *out_similarities,
                                          size_t *out_lens) {
    return guard([&] {
        if (!index || (!queries && query_rows) || !out_indices || !out_similarities || !out_lens)
            throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        if (k == 0) throw Error(NNDEX_ERR_INVALID_TOP_K, "top-k must be greater than zero");
        auto res = index->impl->search_batch(queries, query_rows, k, approx != 0);
        for (size_t q = 0; q < res.size(); ++q) {
            for (size_t i = 0; i < res[q].size(); ++i) {
                out_indices[q * k + i] = res[q][i].index;
                out_similarities[q * k + i] = res[q][i].similarity;
            }
            out_lens[q] = res[q].size();
        }
    });
}

nndex_status nndex_gpu_index_adapter_info(const nndex_gpu_index *index,
                                          nndex_gpu_device_info *out) {
    return guard([&] {
        if (!index || !out) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        *out = index->impl->info();
        out->index = 0;  // NNdex::gpu_info reports index 0 (src/lib.rs:306)
    });
}

nndex_status nndex_list_gpu_devices(nndex_gpu_device_info *out, size_t capacity, size_t *count) {
    return guard([&] {
        if (!count) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        int n = device_count();
        *count = static_cast<size_t>(n);
        for (int i = 0; i < n && static_cast<size_t>(i) < capacity && out; ++i)
            fill_device_info(i, out[i]);
    });
}

nndex_status nndex_workspace_new(const nndex_gpu_index *index, size_t max_queries, size_t max_k,
                                 nndex_workspace **out) {
    return guard([&] {
        if (!index || !out) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        if (max_k == 0) throw Error(NNDEX_ERR_INVALID_TOP_K, "top-k must be greater than zero");
        auto w = new nndex_workspace();
        try {
            w->impl = index->impl->make_workspace(max_queries, max_k);
            w->owner = index->impl.get();
        } catch (...) {
            delete w;
            throw;
        }
        *out = w;
    });
}

void nndex_workspace_free(nndex_workspace *ws) { delete ws; }

nndex_status nndex_gpu_index_search_device(nndex_gpu_index *index, nndex_workspace *ws,
                                           const float *device_queries, size_t query_rows,
                                           size_t k, uint64_t *device_indices,
                                           float *device_similarities, void *stream) {
    return guard([&] {
        if (!index || !ws || !device_queries || !device_indices || !device_similarities)
            throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        if (ws->owner != index->impl.get())
            throw Error(NNDEX_ERR_INVALID_ARGUMENT, "workspace belongs to another index");
        if (k == 0) throw Error(NNDEX_ERR_INVALID_TOP_K, "top-k must be greater than zero");
        index->impl->search_exact_any(*ws->impl, device_queries, query_rows, k,
                                      device_indices, device_similarities,
                                      static_cast<hipStream_t>(stream));
    });
}

nndex_status nndex_gpu_index_search_batch_device(nndex_gpu_index *index, nndex_workspace *ws,
                                                 const float *device_queries, size_t query_rows,
                                                 size_t k, int32_t precision, size_t overfetch,
                                                 uint64_t *device_indices,
                                                 float *device_similarities,
                                                 uint8_t *device_flags, void *stream) {
    return guard([&] {
        if (!index || !ws || !device_queries || !device_indices || !device_similarities ||
            !device_flags)
            throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        if (ws->owner != index->impl.get())
            throw Error(NNDEX_ERR_INVALID_ARGUMENT, "workspace belongs to another index");
        if (k == 0) throw Error(NNDEX_ERR_INVALID_TOP_K, "top-k must be greater than zero");
        if (precision != NNDEX_PRECISION_F32 && precision != NNDEX_PRECISION_BF16 &&
            precision != NNDEX_PRECISION_FP8)
            throw Error(NNDEX_ERR_INVALID_ARGUMENT, "unknown precision");
        index->impl->search_batch_mfma(*ws->impl, device_queries, query_rows, k,
                                       static_cast<Precision>(precision), overfetch,
                                       device_indices, device_similarities, device_flags,
                                       static_cast<hipStream_t>(stream));
    });
}

}  // extern "C"

// ---- Exact search without allocation ---------------------------------------
// The fused scan's scratch lives in this file-scope device array instead of a
// workspace: the block lists (kNoAllocPartKeys 64-bit keys, 8 MiB), then at
// kNoAllocCounters the live counts [8], the thresholds [8] and the ticket,
// each 256-byte aligned.  Every device the library runs on holds its own copy
// until the process ends.  In host code `scratch` names a host-side stand-in:
// the device address comes from hipGetSymbolAddress, once per device.
namespace {
constexpr size_t kNoAllocPartKeys = size_t(1) << 20;
constexpr size_t kNoAllocCounters = kNoAllocPartKeys * sizeof(uint64_t);
constexpr size_t kNoAllocBytes = kNoAllocCounters + 1024;
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kNoAllocMaxDevices] = {};
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Device address of `scratch` on `device` (looked up on the first call there,
// with that device current; the caller's current device is restored).
static unsigned char *scratch_on_device(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices)
        throw Error(NNDEX_ERR_UNSUPPORTED, "device index beyond the no-allocation table");
    std::lock_guard<std::mutex> l(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        NNDEX_HIP_CHECK(hipGetDevice(&cur));
        if (cur != device) NNDEX_HIP_CHECK(hipSetDevice(device));
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess)
            throw Error(NNDEX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                           hipGetErrorString(e));
        if (!ptr) throw Error(NNDEX_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_ptr[device] = ptr;
    }
    return static_cast<unsigned char *>(g_scratch_ptr[device]);
}

extern "C" {

nndex_status nndex_gpu_index_search_device_noalloc(nndex_gpu_index *index,
                                                   const float *device_queries, size_t query_rows,
                                                   size_t k, uint64_t *device_indices,
                                                   float *device_similarities, void *stream) {
    return guard([&] {
        if (!index || !device_indices || !device_similarities ||
            (!device_queries && query_rows))
            throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        if (k == 0) throw Error(NNDEX_ERR_INVALID_TOP_K, "top-k must be greater than zero");
        const GpuIndex &g = *index->impl;
        if (k > g.rows()) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "k must be in [1, rows]");
        if (!g.fused_scan_serves(k))
            throw Error(NNDEX_ERR_UNSUPPORTED,
                        "k=" + std::to_string(k) +
                            " exceeds the fused scan (k <= 992, or k = rows <= 1984)");
        const size_t need = g.scan_part_keys(k);
        if (need > kNoAllocPartKeys)
            throw Error(NNDEX_ERR_UNSUPPORTED,
                        "k=" + std::to_string(k) + " needs " + std::to_string(need) +
                            " block-list keys, above the no-allocation scratch's " +
                            std::to_string(kNoAllocPartKeys) + " (8 MiB)");
        if (query_rows == 0) return;
        unsigned char *base = scratch_on_device(g.device());
        const hipStream_t s = static_cast<hipStream_t>(stream);
        ScanScratch sc;
        sc.part = reinterpret_cast<uint64_t *>(base);
        sc.part_elems = kNoAllocPartKeys;
        sc.part_count = reinterpret_cast<uint32_t *>(base + kNoAllocCounters);
        sc.theta = reinterpret_cast<unsigned long long *>(base + kNoAllocCounters + 256);
        sc.ticket = reinterpret_cast<uint32_t *>(base + kNoAllocCounters + 512);
        // counts, thresholds and ticket start at zero (a kernel, not a
        // memset node, so the call may be captured into a graph)
        launch_zero_u32(reinterpret_cast<uint32_t *>(base + kNoAllocCounters), 1024 / 4, s);
        g.search_device(sc, device_queries, query_rows, k, device_indices, device_similarities,
                        s);
    });
}

}  // extern "C"

struct nndex_graph {
    std::unique_ptr<BatchGraph> impl;
};

extern "C" {

nndex_status nndex_gpu_index_capture_batch(nndex_gpu_index *index, nndex_workspace *ws,
                                           const float *device_queries, size_t query_rows,
                                           size_t k, int32_t precision, size_t overfetch,
                                           uint64_t *device_indices, float *device_similarities,
                                           uint8_t *device_flags, void *stream,
                                           nndex_graph **out) {
    return guard([&] {
        if (!index || !ws || !device_queries || !device_indices || !device_similarities ||
            !device_flags || !out)
            throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        if (!stream) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "capture needs a non-null stream");
        if (ws->owner != index->impl.get())
            throw Error(NNDEX_ERR_INVALID_ARGUMENT, "workspace belongs to another index");
        if (k == 0) throw Error(NNDEX_ERR_INVALID_TOP_K, "top-k must be greater than zero");
        if (precision != NNDEX_PRECISION_F32 && precision != NNDEX_PRECISION_BF16 &&
            precision != NNDEX_PRECISION_FP8)
            throw Error(NNDEX_ERR_INVALID_ARGUMENT, "unknown precision");
        std::unique_ptr<nndex_graph> gr(new nndex_graph());
        gr->impl = index->impl->capture_batch(
            *ws->impl, device_queries, query_rows, k, static_cast<Precision>(precision),
            overfetch, device_indices, device_similarities, device_flags,
            static_cast<hipStream_t>(stream));
        *out = gr.release();
    });
}

nndex_status nndex_graph_launch(nndex_graph *graph, void *stream) {
    return guard([&] {
        if (!graph) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        graph->impl->launch(static_cast<hipStream_t>(stream));
    });
}

void nndex_graph_free(nndex_graph *graph) { delete graph; }

nndex_status nndex_gpu_index_autotune(nndex_gpu_index *index, int32_t reps, char *report,
                                     size_t report_len) {
    return guard([&] {
        if (!index) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        std::string r = index->impl->autotune(reps > 0 ? reps : 20);
        if (report && report_len) {
            size_t n = r.size() < report_len - 1 ? r.size() : report_len - 1;
            std::memcpy(report, r.data(), n);
            report[n] = 0;
        }
    });
}

nndex_status nndex_gpu_index_autotune_batch(nndex_gpu_index *index, int32_t reps, char *report,
                                           size_t report_len) {
    return guard([&] {
        if (!index) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        std::string r = index->impl->autotune_batch(reps > 0 ? reps : 3);
        if (report && report_len) {
            size_t n = r.size() < report_len - 1 ? r.size() : report_len - 1;
            std::memcpy(report, r.data(), n);
            report[n] = 0;
        }
    });
}

nndex_status nndex_gpu_index_set_batch_tiles(nndex_gpu_index *index, int32_t precision,
                                             int32_t variant) {
    return guard([&] {
        if (!index) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        if (precision < 0 || precision > 2 || variant < 0 || variant > 1)
            throw Error(NNDEX_ERR_INVALID_ARGUMENT, "precision must be 0-2 and variant 0-1");
        index->impl->set_batch_variant(static_cast<Precision>(precision), variant);
    });
}

nndex_status nndex_gpu_index_set_launch(nndex_gpu_index *index, uint32_t scan_tiles_per_wave,
                                        uint32_t scan_flags) {
    return guard([&] {
        if (!index) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        index->impl->set_scan_tiles_per_wave(scan_tiles_per_wave);
        index->impl->set_scan_flags(scan_flags);
    });
}

nndex_status nndex_gpu_index_set_ivf_params(nndex_gpu_index *index, size_t nlist, size_t nprobe,
                                            size_t iters) {
    return guard([&] {
        if (!index) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        index->impl->set_ivf_params(nlist, nprobe, iters);
    });
}

nndex_status nndex_gpu_index_build_ivf(nndex_gpu_index *index, nndex_ivf_info *out) {
    return guard([&] {
        if (!index) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        const IvfIndex &ivf = index->impl->build_ivf();
        if (out) {
            out->rows = ivf.rows;
            out->padded_dims = ivf.pdims;
            out->nlist = ivf.nlist;
            out->iters = ivf.iters;
            out->rechecked = ivf.rechecked;
            out->build_ms = ivf.build_ms;
        }
    });
}

nndex_status nndex_gpu_index_ivf_import(nndex_gpu_index *index, size_t nlist,
                                        const float *centroids, const uint64_t *offsets,
                                        const uint64_t *original_indices,
                                        const uint64_t *assignments, nndex_ivf_info *out) {
    return guard([&] {
        if (!index) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        const IvfIndex &ivf =
            index->impl->load_ivf(nlist, centroids, offsets, original_indices, assignments);
        if (out) {
            out->rows = ivf.rows;
            out->padded_dims = ivf.pdims;
            out->nlist = ivf.nlist;
            out->iters = ivf.iters;
            out->rechecked = ivf.rechecked;
            out->build_ms = ivf.build_ms;
        }
    });
}

nndex_status nndex_gpu_index_ivf_export(nndex_gpu_index *index, float *centroids,
                                        uint64_t *offsets, uint64_t *original_indices,
                                        uint64_t *assignments) {
    return guard([&] {
        if (!index) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        const IvfIndex &ivf = index->impl->build_ivf();
        NNDEX_HIP_CHECK(hipSetDevice(index->impl->device()));
        const size_t rows = ivf.rows, nl = ivf.nlist;
        auto widen = [](const uint32_t *d, size_t n, uint64_t *out) {
            std::vector<uint32_t> h(n);
            NNDEX_HIP_CHECK(hipMemcpy(h.data(), d, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; ++i) out[i] = h[i];
        };
        if (nl < 2) {  // single list: identity order (approx.rs:50-60)
            if (offsets) {
                offsets[0] = 0;
                offsets[1] = rows;
            }
            for (size_t i = 0; i < rows; ++i) {
                if (original_indices) original_indices[i] = i;
                if (assignments) assignments[i] = 0;
            }
            return;
        }
        if (centroids)
            NNDEX_HIP_CHECK(hipMemcpy(centroids, ivf.centroids, nl * ivf.pdims * sizeof(float),
                                      hipMemcpyDeviceToHost));
        if (offsets) widen(ivf.offsets, nl + 1, offsets);
        if (original_indices) widen(ivf.orig, rows, original_indices);
        if (assignments) widen(ivf.assign, rows, assignments);
    });
}

nndex_status nndex_gpu_index_search_ivf_device(nndex_gpu_index *index, nndex_workspace *ws,
                                               const float *device_queries, size_t query_rows,
                                               size_t k, size_t nprobe,
                                               uint64_t *device_indices,
                                               float *device_similarities,
                                               uint64_t *device_lens, void *stream) {
    return guard([&] {
        if (!index || !ws || !device_queries || !device_indices || !device_similarities ||
            !device_lens)
            throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        if (ws->owner != index->impl.get())
            throw Error(NNDEX_ERR_INVALID_ARGUMENT, "workspace belongs to another index");
        if (k == 0) throw Error(NNDEX_ERR_INVALID_TOP_K, "top-k must be greater than zero");
        index->impl->search_ivf_device(*ws->impl, device_queries, query_rows, k, nprobe,
                                       device_indices, device_similarities, device_lens,
                                       static_cast<hipStream_t>(stream));
    });
}

nndex_status nndex_device_copy(void *device_dst, const void *device_src, size_t bytes,
                               void *stream) {
    return guard([&] {
        launch_copy(device_dst, device_src, bytes, static_cast<hipStream_t>(stream));
    });
}

nndex_status nndex_debug_mfma_tile(int32_t elem, const void *a, const void *b, uint32_t calls,
                                   float *out) {
    return guard([&] {
        if (!a || !b || !out) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        if (calls == 0 || calls > 4096) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "calls in [1, 4096]");
        const size_t per = elem == kElemF32 ? 4 : (elem == kElemBF16 ? 16 : 32);
        const size_t bytes = per * 64 * calls;
        DeviceBuffer da, db, dout;
        da.reserve(bytes);
        db.reserve(bytes);
        dout.reserve(256 * sizeof(float));
        NNDEX_HIP_CHECK(hipMemcpy(da.ptr, a, bytes, hipMemcpyHostToDevice));
        NNDEX_HIP_CHECK(hipMemcpy(db.ptr, b, bytes, hipMemcpyHostToDevice));
        launch_mfma_tile(elem, da.ptr, db.ptr, calls, static_cast<float *>(dout.ptr), nullptr);
        NNDEX_HIP_CHECK(hipMemcpy(out, dout.ptr, 256 * sizeof(float), hipMemcpyDeviceToHost));
    });
}

nndex_status nndex_device_count(int32_t *count) {
    return guard([&] {
        if (!count) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        *count = device_count();
    });
}
nndex_status nndex_set_device(int32_t device) {
    return guard([&] { NNDEX_HIP_CHECK(hipSetDevice(device)); });
}
nndex_status nndex_device_malloc(void **ptr, size_t bytes) {
    return guard([&] { NNDEX_HIP_CHECK(hipMalloc(ptr, bytes)); });
}
nndex_status nndex_device_free(void *ptr) {
    return guard([&] { NNDEX_HIP_CHECK(hipFree(ptr)); });
}
nndex_status nndex_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    return guard([&] { NNDEX_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}
nndex_status nndex_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    return guard([&] { NNDEX_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}
nndex_status nndex_stream_create(void **stream) {
    return guard([&] {
        hipStream_t s;
        NNDEX_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        *stream = s;
    });
}
nndex_status nndex_stream_destroy(void *stream) {
    return guard([&] { NNDEX_HIP_CHECK(hipStreamDestroy(static_cast<hipStream_t>(stream))); });
}
nndex_status nndex_stream_synchronize(void *stream) {
    return guard(
        [&] { NNDEX_HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream))); });
}
nndex_status nndex_event_create(void **event) {
    return guard([&] {
        hipEvent_t e;
        NNDEX_HIP_CHECK(hipEventCreate(&e));
        *event = e;
    });
}
nndex_status nndex_event_destroy(void *event) {
    return guard([&] { NNDEX_HIP_CHECK(hipEventDestroy(static_cast<hipEvent_t>(event))); });
}
nndex_status nndex_event_record(void *event, void *stream) {
    return guard([&] {
        NNDEX_HIP_CHECK(
            hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)));
    });
}
nndex_status nndex_event_synchronize(void *event) {
    return guard([&] { NNDEX_HIP_CHECK(hipEventSynchronize(static_cast<hipEvent_t>(event))); });
}
nndex_status nndex_event_elapsed_ms(void *start, void *stop, float *ms) {
    return guard([&] {
        NNDEX_HIP_CHECK(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start),
                                            static_cast<hipEvent_t>(stop)));
    });
}

}  // extern "C"
