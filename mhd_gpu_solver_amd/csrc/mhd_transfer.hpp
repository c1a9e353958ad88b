// Host <-> device transfers of a solver's SoA state through pinned staging.
//
// The caller's arrays are ordinary (pageable) fp64 memory. A copy from pageable memory goes
// through the runtime's own bounce buffers at a fraction of the link rate, and the fp32
// storage also needs a conversion. Here the state moves in chunks through two pinned
// buffers: host threads fill (or drain) one buffer, converting between fp64 and the storage
// type on the way, while the copy engine moves the other. Small states keep direct copies.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstddef>
#include <functional>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace mhdhip {

// A few host threads for bulk copies, shared by the process. run() splits [0, n) into one
// contiguous range per thread (the caller takes one) and returns when all are done; calls
// from several threads take turns.
class CopyPool {
public:
    static CopyPool& instance()
    {
        static CopyPool pool;
        return pool;
    }

    void run(size_t n, const std::function<void(size_t, size_t)>& body)
    {
        std::lock_guard<std::mutex> one_job(job_mutex_);
        const size_t parts = workers_.size() + 1;
        const size_t step = (n + parts - 1) / parts;
        {
            std::lock_guard<std::mutex> lk(m_);
            body_ = &body;
            n_ = n;
            step_ = step;
            pending_ = workers_.size();
            ++generation_;
        }
        cv_.notify_all();
        body(std::min(n, (parts - 1) * step), n);  // the caller's share: the last range
        std::unique_lock<std::mutex> lk(m_);
        done_cv_.wait(lk, [&] { return pending_ == 0; });
        body_ = nullptr;
    }

    ~CopyPool()
    {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
            ++generation_;
        }
        cv_.notify_all();
        for (auto& t : workers_) {
            t.join();
        }
    }

private:
    CopyPool()
    {
        const unsigned hw = std::thread::hardware_concurrency();
        const unsigned threads = std::max(1u, std::min(8u, hw == 0 ? 1u : hw));
        for (unsigned k = 0; k + 1 < threads; ++k) {
            workers_.emplace_back([this, k] { work(k); });
        }
    }

    void work(size_t k)
    {
        unsigned long long seen = 0;
        for (;;) {
            const std::function<void(size_t, size_t)>* body;
            size_t lo, hi;
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return generation_ != seen; });
                seen = generation_;
                if (stop_) {
                    return;
                }
                body = body_;
                lo = std::min(n_, k * step_);
                hi = std::min(n_, lo + step_);
            }
            (*body)(lo, hi);
            {
                std::lock_guard<std::mutex> lk(m_);
                --pending_;
            }
            done_cv_.notify_one();
        }
    }

    std::vector<std::thread> workers_;
    std::mutex job_mutex_;
    std::mutex m_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(size_t, size_t)>* body_ = nullptr;
    size_t n_ = 0, step_ = 0, pending_ = 0;
    unsigned long long generation_ = 0;
    bool stop_ = false;
};

// Packs `count` host planes of n elements into one contiguous buffer (gather) or unpacks it
// (scatter), converting the element type; on the copy threads when the state is large.
template <typename Body>
void for_plane_spans(int count, size_t n, Body&& body)
{
    const size_t total = static_cast<size_t>(count) * n;
    auto spans = [&](size_t lo, size_t hi) {
        for (size_t j = lo; j < hi;) {
            const size_t c = j / n;
            const size_t k = j - c * n;
            const size_t m = std::min(hi - j, n - k);
            body(static_cast<int>(c), k, m);
            j += m;
        }
    };
    if (total < (size_t(1) << 20)) {
        spans(0, total);
    } else {
        CopyPool::instance().run(total, spans);
    }
}

template <typename To, typename From>
void gather_planes(To* dst, const From* const* src, int count, size_t n)
{
    for_plane_spans(count, n, [&](int c, size_t k, size_t m) {
        To* d = dst + static_cast<size_t>(c) * n + k;
        const From* f = src[c] + k;
        for (size_t i = 0; i < m; ++i) {
            d[i] = static_cast<To>(f[i]);
        }
    });
}

template <typename To, typename From>
void scatter_planes(To* const* dst, const From* src, int count, size_t n)
{
    for_plane_spans(count, n, [&](int c, size_t k, size_t m) {
        To* d = dst[c] + k;
        const From* f = src + static_cast<size_t>(c) * n + k;
        for (size_t i = 0; i < m; ++i) {
            d[i] = static_cast<To>(f[i]);
        }
    });
}

// Two pinned chunk buffers and their copy-done events, owned by one solver.
class StagedCopier {
public:
    static constexpr size_t kChunkBytes = size_t(16) << 20;
    // below this many state bytes the direct copies are as fast
    static constexpr size_t kMinBytes = size_t(8) << 20;

    ~StagedCopier() { release(); }

    void release()
    {
        for (int b = 0; b < 2; ++b) {
            if (buf_[b]) (void)hipHostFree(buf_[b]);
            if (done_[b]) (void)hipEventDestroy(done_[b]);
            buf_[b] = nullptr;
            done_[b] = nullptr;
        }
    }

    // count planes of n elements: dst + c * plane <- (R) comps[c][0..n)
    template <typename R>
    void upload(R* dst, size_t plane, const double* const* comps, int count, size_t n,
                hipStream_t s)
    {
        ensure();
        const size_t chunk = kChunkBytes / sizeof(R);
        size_t k = 0;
        for (int c = 0; c < count; ++c) {
            for (size_t o = 0; o < n; o += chunk, ++k) {
                const int b = static_cast<int>(k & 1);
                const size_t m = std::min(chunk, n - o);
                if (k >= 2) {
                    check(hipEventSynchronize(done_[b]), "hipEventSynchronize(staging)");
                }
                R* stage = static_cast<R*>(buf_[b]);
                const double* from = comps[c] + o;
                CopyPool::instance().run(m, [&](size_t lo, size_t hi) {
                    for (size_t i = lo; i < hi; ++i) {
                        stage[i] = static_cast<R>(from[i]);
                    }
                });
                check(hipMemcpyAsync(dst + c * plane + o, stage, m * sizeof(R),
                                     hipMemcpyHostToDevice, s),
                      "hipMemcpyAsync(upload)");
                check(hipEventRecord(done_[b], s), "hipEventRecord(staging)");
            }
        }
        check(hipStreamSynchronize(s), "hipStreamSynchronize(upload)");
    }

    // comps[c][0..n) <- (double) (src + c * plane), for count planes
    template <typename R>
    void download(double* const* comps, const R* src, size_t plane, int count, size_t n,
                  hipStream_t s)
    {
        ensure();
        const size_t chunk = kChunkBytes / sizeof(R);
        const size_t per_plane = (n + chunk - 1) / chunk;
        const size_t total = per_plane * static_cast<size_t>(count);
        auto span = [&](size_t k, int& c, size_t& o, size_t& m) {
            c = static_cast<int>(k / per_plane);
            o = (k % per_plane) * chunk;
            m = std::min(chunk, n - o);
        };
        auto issue = [&](size_t k) {
            int c;
            size_t o, m;
            span(k, c, o, m);
            const int b = static_cast<int>(k & 1);
            check(hipMemcpyAsync(buf_[b], src + c * plane + o, m * sizeof(R),
                                 hipMemcpyDeviceToHost, s),
                  "hipMemcpyAsync(download)");
            check(hipEventRecord(done_[b], s), "hipEventRecord(staging)");
        };
        for (size_t k = 0; k < std::min<size_t>(2, total); ++k) {
            issue(k);
        }
        for (size_t k = 0; k < total; ++k) {
            int c;
            size_t o, m;
            span(k, c, o, m);
            const int b = static_cast<int>(k & 1);
            check(hipEventSynchronize(done_[b]), "hipEventSynchronize(staging)");
            const R* stage = static_cast<const R*>(buf_[b]);
            double* to = comps[c] + o;
            CopyPool::instance().run(m, [&](size_t lo, size_t hi) {
                for (size_t i = lo; i < hi; ++i) {
                    to[i] = static_cast<double>(stage[i]);
                }
            });
            if (k + 2 < total) {
                issue(k + 2);
            }
        }
    }

private:
    static void check(hipError_t e, const char* what)
    {
        if (e != hipSuccess) {
            throw std::runtime_error(std::string("HIP error in ") + what + ": " +
                                     hipGetErrorString(e));
        }
    }

    void ensure()
    {
        for (int b = 0; b < 2; ++b) {
            if (!buf_[b]) {
                check(hipHostMalloc(&buf_[b], kChunkBytes, 0), "hipHostMalloc(staging)");
            }
            if (!done_[b]) {
                check(hipEventCreateWithFlags(&done_[b], hipEventDisableTiming),
                      "hipEventCreate(staging)");
            }
        }
    }

    void* buf_[2] = {nullptr, nullptr};
    hipEvent_t done_[2] = {nullptr, nullptr};
};

}  // namespace mhdhip
