// Serving loop: continuous batching of search requests over one resident
// index.  No reference counterpart (the reference serves one call at a time
// behind a Mutex, src/gpu.rs:511-592); its searches are the ones this loop
// batches: exact (src/gpu.rs:486-785) and IVF (src/gpu.rs:435-462).
//
// Structure:
//   * Requests (query rows, k, mode exact | IVF, nprobe) enter a FIFO.
//   * `slots` executor threads each own a HIP stream, a workspace prepared for
//     the largest batch, device query/result buffers and pinned host staging.
//     A free executor takes the oldest pending request and every later pending
//     request of the same kind (mode, nprobe) that still fits, up to
//     max_batch_queries rows; a request larger than that is split across
//     batches.  The batch runs at the largest k of its requests; each request
//     receives the first k of its rows' results (the exact and IVF orders are
//     total, so the first k of a top-k' list are the top-k).
//   * Work that arrives while batches run is admitted as soon as an executor
//     frees, so batches of different kinds overlap on the device.
//   * With use_graphs, exact batches above 4 rows replay a HIP graph captured
//     per executor for (rows rounded up to a multiple of 16 up to 128 rows and
//     of 64 above, k rounded up to a power of two):
//     padding rows repeat the last row, each row keeps its own first k
//     results, and rows the certificate leaves open run the exact fused scan.
//   * Admission by HBM footprint: an executor's device memory is fixed at
//     start (workspace for max_batch_queries x max_k of each enabled mode plus
//     query/result buffers); the number of executors is the largest that fits
//     the byte budget, and no allocation happens while serving.
// Latency is measured from submit to the moment a request's last rows are
// copied out; p50/p99 are over completed requests.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gpu_index.hpp"

namespace nndex {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t id = 0;
    int mode = 0;          // 0 exact, 1 IVF
    uint32_t nprobe = 0;   // IVF: resolved probe count
    uint32_t k = 0;        // results per row (capped at rows)
    uint32_t nq = 0;
    // raw requests (raw_dims > 0): rows of raw_dims values, checked and
    // prepared on the device in raw_mode (kQueryPad / kQueryBatch)
    uint32_t raw_dims = 0;
    int raw_mode = 0;
    std::vector<float> q;  // nq x (raw_dims ? raw_dims : pdims)
    std::vector<uint64_t> idx, lens;
    std::vector<float> sim;
    uint32_t issued = 0, done = 0;
    Clock::time_point t_submit, t_done;
    nndex_status status = NNDEX_OK;
    std::string error;
    bool finished = false;
};

struct Segment {
    std::shared_ptr<Request> req;
    uint32_t q0, n;
    uint32_t row_flags = 0;  // raw rows: OR of the device checks (1 non-finite, 2 zero norm)
};

struct Executor {
    hipStream_t stream = nullptr;
    std::unique_ptr<Workspace> ws;
    DeviceBuffer dq, didx, dsim, dlen, dflags, draw, drowflags, dflagword;
    PinnedBuffer hq, hidx, hsim, hlen, hflags, hrowflags;
    // captured exact batches by (rows bucket, k bucket), made on first use
    std::map<std::pair<uint32_t, uint32_t>, std::unique_ptr<BatchGraph>> graphs;
};

class Server {
  public:
    Server(std::shared_ptr<GpuIndex> index, const nndex_server_options &o);
    ~Server();
    // raw_dims == 0: normalised padded rows; else rows of raw_dims values
    // prepared on the device (normalize: normalize_rows, else ensure_finite +
    // pad), a failed check reported by wait for that request
    uint64_t submit(const float *q, size_t nq, size_t k, int mode, size_t nprobe,
                    size_t raw_dims = 0, bool normalize = false);
    void wait(uint64_t ticket, uint64_t *idx, float *sim, uint64_t *lens, double *latency_ms);
    void stats(nndex_server_stats &out);
    void reset_stats();

  private:
    void run(size_t e);
    bool next_batch(std::vector<Segment> &batch);
    // runs the batch; returns true when it replayed a captured graph
    bool execute(Executor &ex, std::vector<Segment> &batch);

    std::shared_ptr<GpuIndex> index_;
    nndex_server_options opt_;
    size_t pdims_ = 0, rows_ = 0;
    std::vector<std::unique_ptr<Executor>> ex_;
    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<std::shared_ptr<Request>> pending_;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> live_;
    uint64_t next_id_ = 1;
    bool stop_ = false;
    size_t slot_bytes_ = 0;
    // statistics since the last reset
    std::vector<double> lat_ms_;
    uint64_t s_requests_ = 0, s_queries_ = 0, s_batches_ = 0, s_batch_rows_ = 0,
             s_graph_batches_ = 0;
    Clock::time_point s_first_, s_last_;
    bool s_any_ = false;
};

Server::Server(std::shared_ptr<GpuIndex> index, const nndex_server_options &o)
    : index_(std::move(index)), opt_(o) {
    if (opt_.slots == 0) opt_.slots = 4;
    if (opt_.max_batch_queries == 0) opt_.max_batch_queries = 2048;
    if (opt_.max_k == 0) opt_.max_k = 100;
    if (opt_.hbm_budget_bytes == 0) opt_.hbm_budget_bytes = 16ull << 30;
    pdims_ = index_->padded_dims();
    rows_ = index_->rows();
    if (opt_.max_k > rows_) opt_.max_k = static_cast<uint32_t>(rows_);
    const size_t B = opt_.max_batch_queries, K = opt_.max_k;
    const size_t np = opt_.ivf_nprobe_max;
    NNDEX_HIP_CHECK(hipSetDevice(index_->device()));
    // footprint of one executor: workspace + device query/result buffers
    slot_bytes_ = index_->workspace_bytes(B, K, np) + 2 * B * pdims_ * 4 + B * K * 12 + B * 10 +
                  256;
    size_t slots = opt_.hbm_budget_bytes / slot_bytes_;
    if (slots > opt_.slots) slots = opt_.slots;
    if (slots == 0)
        throw Error(NNDEX_ERR_OUT_OF_MEMORY,
                    "HBM budget of " + std::to_string(opt_.hbm_budget_bytes) +
                        " bytes is below one executor's footprint of " +
                        std::to_string(slot_bytes_) + " bytes");
    opt_.slots = static_cast<uint32_t>(slots);
    for (size_t e = 0; e < slots; ++e) {
        std::unique_ptr<Executor> ex(new Executor());
        NNDEX_HIP_CHECK(hipStreamCreateWithFlags(&ex->stream, hipStreamNonBlocking));
        ex->ws = index_->make_workspace(8, K < 10 ? 10 : K);
        index_->prepare(*ex->ws, B, K, np, ex->stream);
        ex->dq.reserve(B * pdims_ * 4);
        ex->didx.reserve(B * K * 8);
        ex->dsim.reserve(B * K * 4);
        ex->dlen.reserve(B * 8);
        ex->dflags.reserve(B);
        ex->hflags.reserve(B);
        ex->draw.reserve(B * pdims_ * 4);
        ex->drowflags.reserve(B);
        ex->dflagword.reserve(sizeof(uint32_t));
        ex->hrowflags.reserve(B);
        ex->hq.reserve(B * pdims_ * 4);
        ex->hidx.reserve(B * K * 8);
        ex->hsim.reserve(B * K * 4);
        ex->hlen.reserve(B * 8);
        ex_.push_back(std::move(ex));
    }
    for (size_t e = 0; e < ex_.size(); ++e) threads_.emplace_back(&Server::run, this, e);
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> l(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    for (auto &t : threads_) t.join();
    for (auto &ex : ex_)
        if (ex->stream) {
            (void)hipStreamSynchronize(ex->stream);
            (void)hipStreamDestroy(ex->stream);
        }
}

uint64_t Server::submit(const float *q, size_t nq, size_t k, int mode, size_t nprobe,
                        size_t raw_dims, bool normalize) {
    if (nq == 0) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "a request needs at least one row");
    if (raw_dims && (raw_dims + 7) / 8 * 8 != pdims_)
        throw Error(NNDEX_ERR_QUERY_DIM_MISMATCH,
                    "query dimensionality mismatch: rows of " + std::to_string(raw_dims) +
                        " values do not pad to the index's " + std::to_string(pdims_));
    if (k == 0) throw Error(NNDEX_ERR_INVALID_TOP_K, "top-k must be greater than zero");
    if (mode != 0 && mode != 1) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "mode must be 0 or 1");
    const size_t kc = k < rows_ ? k : rows_;
    if (kc > opt_.max_k)
        throw Error(NNDEX_ERR_INVALID_ARGUMENT,
                    "k above the server's max_k (" + std::to_string(opt_.max_k) + ")");
    auto r = std::make_shared<Request>();
    r->mode = mode;
    r->k = static_cast<uint32_t>(kc);
    r->nq = static_cast<uint32_t>(nq);
    if (mode == 1) {
        if (!opt_.ivf_nprobe_max)
            throw Error(NNDEX_ERR_INVALID_ARGUMENT, "the server was created without IVF");
        const IvfIndex &ivf = index_->build_ivf();
        if (ivf.nlist < 2)
            throw Error(NNDEX_ERR_UNSUPPORTED, "the IVF index has a single list (rows < 256)");
        size_t np = nprobe ? nprobe : ivf_choose_nprobe(ivf.nlist, kc, rows_);
        if (np > ivf.nlist) np = ivf.nlist;
        if (np > opt_.ivf_nprobe_max)
            throw Error(NNDEX_ERR_INVALID_ARGUMENT,
                        "nprobe above the server's ivf_nprobe_max (" +
                            std::to_string(opt_.ivf_nprobe_max) + ")");
        r->nprobe = static_cast<uint32_t>(np);
    }
    r->raw_dims = static_cast<uint32_t>(raw_dims);
    r->raw_mode = normalize ? kQueryBatch : kQueryPad;
    r->q.assign(q, q + nq * (raw_dims ? raw_dims : pdims_));
    r->idx.assign(nq * kc, 0);
    r->sim.assign(nq * kc, 0.f);
    r->lens.assign(nq, 0);
    uint64_t id;
    {
        std::lock_guard<std::mutex> l(mu_);
        id = r->id = next_id_++;
        r->t_submit = Clock::now();
        pending_.push_back(r);
        live_[id] = r;
    }
    cv_work_.notify_one();
    return id;
}

// Under mu_: the oldest pending request and every later one of the same kind
// that fits, up to max_batch_queries rows.
bool Server::next_batch(std::vector<Segment> &batch) {
    batch.clear();
    while (!pending_.empty() && pending_.front()->issued == pending_.front()->nq)
        pending_.pop_front();
    if (pending_.empty()) return false;
    const int mode = pending_.front()->mode;
    const uint32_t nprobe = pending_.front()->nprobe;
    const uint32_t raw_dims = pending_.front()->raw_dims;
    const int raw_mode = pending_.front()->raw_mode;
    uint32_t rows = 0;
    for (auto it = pending_.begin(); it != pending_.end() && rows < opt_.max_batch_queries;) {
        Request &r = **it;
        if (r.mode != mode || r.nprobe != nprobe || r.raw_dims != raw_dims ||
            (raw_dims && r.raw_mode != raw_mode) || r.issued == r.nq) {
            ++it;
            continue;
        }
        const uint32_t n = std::min(r.nq - r.issued, opt_.max_batch_queries - rows);
        batch.push_back({*it, r.issued, n});
        r.issued += n;
        rows += n;
        if (r.issued == r.nq)
            it = pending_.erase(it);
        else
            ++it;
    }
    return true;
}

bool Server::execute(Executor &ex, std::vector<Segment> &batch) {
    uint32_t nq = 0, k = 0;
    for (auto &sg : batch) {
        nq += sg.n;
        k = std::max(k, sg.req->k);
    }
    const int mode = batch[0].req->mode;
    const uint32_t nprobe = batch[0].req->nprobe;
    const uint32_t raw_dims = batch[0].req->raw_dims;
    const int raw_mode = batch[0].req->raw_mode;
    const size_t cols = raw_dims ? raw_dims : pdims_;  // staged values per row
    float *hq = static_cast<float *>(ex.hq.ptr);
    size_t off = 0;
    for (auto &sg : batch) {
        std::memcpy(hq + off * cols, sg.req->q.data() + (size_t)sg.q0 * cols,
                    (size_t)sg.n * cols * sizeof(float));
        off += sg.n;
    }
    // rows to the device: as they are, or raw rows checked and prepared into
    // dq there (their per-row flags come back with the results)
    auto upload = [&](uint32_t rows) {
        if (!raw_dims) {
            NNDEX_HIP_CHECK(hipMemcpyAsync(ex.dq.ptr, hq, (size_t)rows * pdims_ * 4,
                                           hipMemcpyHostToDevice, ex.stream));
            return;
        }
        NNDEX_HIP_CHECK(hipMemcpyAsync(ex.draw.ptr, hq, (size_t)rows * cols * 4,
                                       hipMemcpyHostToDevice, ex.stream));
        launch_normalize_rowmajor(static_cast<const float *>(ex.draw.ptr), rows, raw_dims,
                                  static_cast<uint32_t>(pdims_), raw_mode,
                                  static_cast<float *>(ex.dq.ptr),
                                  static_cast<uint32_t *>(ex.dflagword.ptr),
                                  static_cast<uint8_t *>(ex.drowflags.ptr), ex.stream);
        NNDEX_HIP_CHECK(hipMemcpyAsync(ex.hrowflags.ptr, ex.drowflags.ptr, rows,
                                       hipMemcpyDeviceToHost, ex.stream));
    };
    const float *dq = static_cast<const float *>(ex.dq.ptr);
    uint64_t *didx = static_cast<uint64_t *>(ex.didx.ptr);
    float *dsim = static_cast<float *>(ex.dsim.ptr);
    uint32_t kstride = k;  // result row stride of this batch
    bool graphed = false;
    Precision prec;
    if (mode == 0 && opt_.use_graphs && nq > kFusedBatchMax &&
        index_->exact_batch_precision(prec)) {
        // replay a captured exact batch: rows padded (with copies of the
        // last row) to a multiple of 16 up to 128 rows and of 64 above, so
        // at most one query chunk of the scan is padding; k to a power of
        // two (results are prefixes)
        uint32_t qb = nq <= 128 ? (nq + 15) / 16 * 16 : (nq + 63) / 64 * 64, kb = 1;
        if (qb > opt_.max_batch_queries) qb = opt_.max_batch_queries;
        while (kb < k) kb <<= 1;
        if (kb > opt_.max_k) kb = opt_.max_k;
        for (uint32_t r = nq; r < qb; ++r)
            std::memcpy(hq + (size_t)r * cols, hq + (size_t)(nq - 1) * cols, cols * 4);
        upload(qb);
        auto &g = ex.graphs[{qb, kb}];
        uint8_t *dflags = static_cast<uint8_t *>(ex.dflags.ptr);
        // (re)capture when new, or when a workspace buffer moved since
        if (!g || g->generation != ex.ws->generation)
            g = index_->capture_batch(*ex.ws, dq, qb, kb, prec, 0, didx, dsim, dflags, ex.stream);
        g->launch(ex.stream);
        NNDEX_HIP_CHECK(hipMemcpyAsync(ex.hflags.ptr, dflags, qb, hipMemcpyDeviceToHost,
                                       ex.stream));
        NNDEX_HIP_CHECK(hipStreamSynchronize(ex.stream));
        // rows the certificate left open: the exact fused scan, as in the
        // exact batch search
        const uint8_t *hf = static_cast<const uint8_t *>(ex.hflags.ptr);
        for (uint32_t q = 0; q < nq; ++q)
            if (hf[q])
                index_->search_device(*ex.ws, dq + (size_t)q * pdims_, 1, kb,
                                      didx + (size_t)q * kb, dsim + (size_t)q * kb, ex.stream);
        kstride = kb;
        graphed = true;
    } else {
        upload(nq);
        if (mode == 0)
            index_->search_exact_any(*ex.ws, dq, nq, k, didx, dsim, ex.stream);
        else
            index_->search_ivf_device(*ex.ws, dq, nq, k, nprobe, didx, dsim,
                                      static_cast<uint64_t *>(ex.dlen.ptr), ex.stream);
    }
    NNDEX_HIP_CHECK(hipMemcpyAsync(ex.hidx.ptr, didx, (size_t)nq * kstride * 8,
                                   hipMemcpyDeviceToHost, ex.stream));
    NNDEX_HIP_CHECK(hipMemcpyAsync(ex.hsim.ptr, dsim, (size_t)nq * kstride * 4,
                                   hipMemcpyDeviceToHost, ex.stream));
    if (mode == 1)
        NNDEX_HIP_CHECK(hipMemcpyAsync(ex.hlen.ptr, ex.dlen.ptr, (size_t)nq * 8,
                                       hipMemcpyDeviceToHost, ex.stream));
    NNDEX_HIP_CHECK(hipStreamSynchronize(ex.stream));
    uint64_t *hi = static_cast<uint64_t *>(ex.hidx.ptr);
    float *hs = static_cast<float *>(ex.hsim.ptr);
    const uint64_t *hl = static_cast<const uint64_t *>(ex.hlen.ptr);
    // A request with k below the batch's k keeps the first k results, unless
    // rows tie across its k-th score: the reference heap's choice among them
    // depends on k, so such a query is searched again alone.
    {
        bool again = false;
        off = 0;
        for (auto &sg : batch) {
            const uint32_t rk = sg.req->k;
            for (uint32_t i = 0; i < sg.n; ++i) {
                const size_t src = (off + i) * kstride;
                if (rk < 2 || rk >= kstride ||
                    __builtin_bit_cast(uint32_t, hs[src + rk - 1]) !=
                        __builtin_bit_cast(uint32_t, hs[src + rk]))
                    continue;
                if (mode == 0)
                    index_->search_device(*ex.ws, dq + (off + i) * pdims_, 1, rk, didx + src,
                                          dsim + src, ex.stream);
                else
                    index_->search_ivf_device(*ex.ws, dq + (off + i) * pdims_, 1, rk, nprobe,
                                              didx + src, dsim + src,
                                              static_cast<uint64_t *>(ex.dlen.ptr) + off + i,
                                              ex.stream);
                NNDEX_HIP_CHECK(hipMemcpyAsync(hi + src, didx + src, rk * 8,
                                               hipMemcpyDeviceToHost, ex.stream));
                NNDEX_HIP_CHECK(hipMemcpyAsync(hs + src, dsim + src, rk * 4,
                                               hipMemcpyDeviceToHost, ex.stream));
                again = true;
            }
            off += sg.n;
        }
        if (again && mode == 1)
            NNDEX_HIP_CHECK(hipMemcpyAsync(ex.hlen.ptr, ex.dlen.ptr, (size_t)nq * 8,
                                           hipMemcpyDeviceToHost, ex.stream));
        if (again) NNDEX_HIP_CHECK(hipStreamSynchronize(ex.stream));
    }
    // raw rows that failed their checks fail their own request only (every
    // row's search is independent of its batch mates)
    if (raw_dims) {
        const uint8_t *rf = static_cast<const uint8_t *>(ex.hrowflags.ptr);
        off = 0;
        for (auto &sg : batch) {
            for (uint32_t i = 0; i < sg.n; ++i) sg.row_flags |= rf[off + i];
            off += sg.n;
        }
    }
    // scatter: each request keeps the first k of its rows' results
    off = 0;
    for (auto &sg : batch) {
        Request &r = *sg.req;
        for (uint32_t i = 0; i < sg.n; ++i) {
            const size_t src = (off + i) * kstride, dst = (size_t)(sg.q0 + i) * r.k;
            std::memcpy(&r.idx[dst], hi + src, r.k * 8);
            std::memcpy(&r.sim[dst], hs + src, r.k * 4);
            const uint64_t found = mode == 1 ? hl[off + i] : kstride;
            r.lens[sg.q0 + i] = found < r.k ? found : r.k;
        }
        off += sg.n;
    }
    return graphed;
}

void Server::run(size_t e) {
    Executor &ex = *ex_[e];
    (void)hipSetDevice(index_->device());
    std::vector<Segment> batch;
    for (;;) {
        {
            std::unique_lock<std::mutex> l(mu_);
            cv_work_.wait(l, [&] {
                if (stop_) return true;
                for (auto &r : pending_)
                    if (r->issued < r->nq) return true;
                return false;
            });
            if (stop_) return;
            if (!next_batch(batch)) continue;
        }
        nndex_status st = NNDEX_OK;
        std::string err;
        bool graphed = false;
        try {
            graphed = execute(ex, batch);
        } catch (const Error &x) {
            st = x.code;
            err = x.what();
        } catch (const std::exception &x) {
            st = NNDEX_ERR_HIP;
            err = x.what();
        }
        uint32_t rows = 0;
        {
            std::lock_guard<std::mutex> l(mu_);
            const auto now = Clock::now();
            for (auto &sg : batch) {
                Request &r = *sg.req;
                rows += sg.n;
                if (st != NNDEX_OK && r.status == NNDEX_OK) {
                    r.status = st;
                    r.error = err;
                }
                if (st == NNDEX_OK && sg.row_flags && r.status == NNDEX_OK) {
                    try {
                        throw_flags(sg.row_flags);
                    } catch (const Error &x) {
                        r.status = x.code;
                        r.error = x.what();
                    }
                }
                r.done += sg.n;
                if (r.done == r.nq) {
                    r.t_done = now;
                    r.finished = true;
                    if (r.status == NNDEX_OK) {
                        lat_ms_.push_back(
                            std::chrono::duration<double, std::milli>(r.t_done - r.t_submit)
                                .count());
                        ++s_requests_;
                        s_queries_ += r.nq;
                        if (!s_any_ || r.t_submit < s_first_) s_first_ = r.t_submit;
                        s_last_ = now;
                        s_any_ = true;
                    }
                }
            }
            ++s_batches_;
            s_batch_rows_ += rows;
            if (graphed) ++s_graph_batches_;
        }
        cv_done_.notify_all();
        if (st != NNDEX_OK) cv_work_.notify_all();
    }
}

void Server::wait(uint64_t ticket, uint64_t *idx, float *sim, uint64_t *lens, double *latency_ms) {
    std::shared_ptr<Request> r;
    {
        std::unique_lock<std::mutex> l(mu_);
        auto it = live_.find(ticket);
        if (it == live_.end())
            throw Error(NNDEX_ERR_INVALID_ARGUMENT,
                        "unknown or already collected ticket " + std::to_string(ticket));
        r = it->second;
        cv_done_.wait(l, [&] { return r->finished; });
        live_.erase(it);
    }
    if (r->status != NNDEX_OK) throw Error(r->status, r->error);
    if (idx) std::memcpy(idx, r->idx.data(), r->idx.size() * 8);
    if (sim) std::memcpy(sim, r->sim.data(), r->sim.size() * 4);
    if (lens) std::memcpy(lens, r->lens.data(), r->lens.size() * 8);
    if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
}

static double percentile(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    // nearest-rank percentile
    size_t rank = static_cast<size_t>(std::ceil(p / 100.0 * v.size()));
    if (rank < 1) rank = 1;
    return v[rank - 1];
}

void Server::stats(nndex_server_stats &out) {
    std::lock_guard<std::mutex> l(mu_);
    std::memset(&out, 0, sizeof(out));
    out.slots = ex_.size();
    out.slot_bytes = slot_bytes_;
    out.requests = s_requests_;
    out.queries = s_queries_;
    out.batches = s_batches_;
    out.graph_batches = s_graph_batches_;
    out.mean_batch_queries = s_batches_ ? (double)s_batch_rows_ / s_batches_ : 0.0;
    out.p50_ms = percentile(lat_ms_, 50);
    out.p99_ms = percentile(lat_ms_, 99);
    out.max_ms = lat_ms_.empty() ? 0.0 : *std::max_element(lat_ms_.begin(), lat_ms_.end());
    const double span =
        s_any_ ? std::chrono::duration<double>(s_last_ - s_first_).count() : 0.0;
    out.span_s = span;
    out.queries_per_s = span > 0 ? s_queries_ / span : 0.0;
    out.requests_per_s = span > 0 ? s_requests_ / span : 0.0;
}

void Server::reset_stats() {
    std::lock_guard<std::mutex> l(mu_);
    lat_ms_.clear();
    s_requests_ = s_queries_ = s_batches_ = s_batch_rows_ = s_graph_batches_ = 0;
    s_any_ = false;
}

}  // namespace nndex

using namespace nndex;

struct nndex_server {
    std::unique_ptr<Server> impl;
};

extern "C" {

nndex_status nndex_server_new(nndex_gpu_index *index, const nndex_server_options *options,
                              nndex_server **out) {
    return guard([&] {
        if (!index || !out) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        nndex_server_options o{};
        if (options) o = *options;
        std::unique_ptr<nndex_server> s(new nndex_server());
        s->impl.reset(new Server(index->impl, o));
        *out = s.release();
    });
}

void nndex_server_free(nndex_server *server) { delete server; }

nndex_status nndex_server_submit(nndex_server *server, const float *queries, size_t query_rows,
                                 size_t k, int32_t mode, size_t nprobe, uint64_t *ticket) {
    return guard([&] {
        if (!server || !queries || !ticket) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        *ticket = server->impl->submit(queries, query_rows, k, mode, nprobe);
    });
}

nndex_status nndex_server_submit_raw(nndex_server *server, const float *queries,
                                     size_t query_rows, size_t dims, int32_t normalize, size_t k,
                                     int32_t mode, size_t nprobe, uint64_t *ticket) {
    return guard([&] {
        if (!server || !queries || !ticket) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        if (dims == 0) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "rows need at least one value");
        *ticket = server->impl->submit(queries, query_rows, k, mode, nprobe, dims, normalize != 0);
    });
}

nndex_status nndex_server_wait(nndex_server *server, uint64_t ticket, uint64_t *out_indices,
                               float *out_similarities, uint64_t *out_lens, double *latency_ms) {
    return guard([&] {
        if (!server) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        server->impl->wait(ticket, out_indices, out_similarities, out_lens, latency_ms);
    });
}

nndex_status nndex_server_get_stats(nndex_server *server, nndex_server_stats *out) {
    return guard([&] {
        if (!server || !out) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        server->impl->stats(*out);
    });
}

nndex_status nndex_server_reset_stats(nndex_server *server) {
    return guard([&] {
        if (!server) throw Error(NNDEX_ERR_INVALID_ARGUMENT, "null pointer");
        server->impl->reset_stats();
    });
}

}  // extern "C"
