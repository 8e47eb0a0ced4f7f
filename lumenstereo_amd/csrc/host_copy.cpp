// host_copy.cpp - a persistent pool of copy threads (see host_copy.h).
#include "host_copy.h"

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>

namespace lumen {

namespace {

constexpr size_t kChunk = 1 << 20;           // bytes per unit of work
constexpr size_t kParallelMin = 4u << 20;    // below this, the caller copies alone
constexpr unsigned kMaxWorkers = 7;          // plus the caller

// A unit of work: `bytes` bytes copied, or, when rows > 0, `rows` rows of a BGR -> gray
// conversion (W pixels a row, source rows srcStride bytes apart, destination rows packed).
struct Chunk {
    char* dst;
    const char* src;
    size_t bytes;
    int rows = 0;
    int W = 0;
    size_t srcStride = 0;
};

void bgr_rows(const Chunk& c) {
    const unsigned char* s0 = reinterpret_cast<const unsigned char*>(c.src);
    unsigned char* d0 = reinterpret_cast<unsigned char*>(c.dst);
    for (int y = 0; y < c.rows; ++y) {
        const unsigned char* s = s0 + (size_t)y * c.srcStride;
        unsigned char* d = d0 + (size_t)y * c.W;
        for (int x = 0; x < c.W; ++x) {
            const int b = s[3 * x], g = s[3 * x + 1], r = s[3 * x + 2];
            d[x] = static_cast<unsigned char>((b * 1868 + g * 9617 + r * 4899 + (1 << 13)) >> 14);
        }
    }
}

void run_chunk(const Chunk& c) {
    if (c.rows > 0)
        bgr_rows(c);
    else
        std::memcpy(c.dst, c.src, c.bytes);
}

class Pool {
public:
    static Pool& get() {
        static Pool pool;
        return pool;
    }

    void run(const std::vector<Chunk>& chunks) {
        std::lock_guard<std::mutex> serial(runMu_);  // one parallel copy at a time
        {
            std::lock_guard<std::mutex> g(mu_);
            work_ = &chunks;
            next_.store(0);
            pending_ = workers_.size();
            ++generation_;
        }
        cv_.notify_all();
        drain(chunks);
        std::unique_lock<std::mutex> lk(mu_);
        doneCv_.wait(lk, [&] { return pending_ == 0; });
        work_ = nullptr;
    }

    ~Pool() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
            ++generation_;
        }
        cv_.notify_all();
        for (std::thread& t : workers_) t.join();
    }

private:
    Pool() {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const unsigned n = std::min(kMaxWorkers, hw > 1 ? hw - 1 : 0u);
        for (unsigned i = 0; i < n; ++i) workers_.emplace_back([this] { loop(); });
    }

    void drain(const std::vector<Chunk>& chunks) {
        for (;;) {
            const size_t i = next_.fetch_add(1);
            if (i >= chunks.size()) return;
            run_chunk(chunks[i]);
        }
    }

    void loop() {
        unsigned long seen = 0;
        for (;;) {
            const std::vector<Chunk>* work;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || generation_ != seen; });
                if (stop_) return;
                seen = generation_;
                work = work_;
            }
            if (work) drain(*work);
            std::lock_guard<std::mutex> g(mu_);
            if (--pending_ == 0) doneCv_.notify_all();
        }
    }

    std::mutex runMu_, mu_;
    std::condition_variable cv_, doneCv_;
    std::vector<std::thread> workers_;
    const std::vector<Chunk>* work_ = nullptr;
    std::atomic<size_t> next_{0};
    size_t pending_ = 0;
    unsigned long generation_ = 0;
    bool stop_ = false;
};

void split(std::vector<Chunk>& out, void* dst, const void* src, size_t bytes) {
    char* d = static_cast<char*>(dst);
    const char* s = static_cast<const char*>(src);
    for (size_t off = 0; off < bytes; off += kChunk) {
        Chunk c{};
        c.dst = d + off;
        c.src = s + off;
        c.bytes = std::min(kChunk, bytes - off);
        out.push_back(c);
    }
}

}  // namespace

void parallel_copy(const std::vector<CopyJob>& jobs) {
    size_t total = 0;
    for (const CopyJob& j : jobs) total += j.bytes;
    if (total < kParallelMin) {
        for (const CopyJob& j : jobs)
            if (j.bytes) std::memcpy(j.dst, j.src, j.bytes);
        return;
    }
    std::vector<Chunk> chunks;
    chunks.reserve(total / kChunk + jobs.size());
    for (const CopyJob& j : jobs) split(chunks, j.dst, j.src, j.bytes);
    Pool::get().run(chunks);
}

void parallel_copy_rows(void* dst, size_t dstStride, const void* src, size_t srcStride,
                        size_t rowBytes, int rows) {
    if (rows <= 0 || rowBytes == 0) return;
    if (dstStride == rowBytes && srcStride == rowBytes) {
        parallel_copy({{dst, src, rowBytes * rows}});
        return;
    }
    std::vector<CopyJob> jobs;
    jobs.reserve(rows);
    for (int y = 0; y < rows; ++y)
        jobs.push_back({static_cast<char*>(dst) + (size_t)y * dstStride,
                        static_cast<const char*>(src) + (size_t)y * srcStride, rowBytes});
    parallel_copy(jobs);
}

void parallel_bgr_to_gray(unsigned char* dst, const unsigned char* src, size_t srcStride, int W,
                          int H) {
    if (W <= 0 || H <= 0) return;
    // Bands of about one chunk of source bytes.
    const int band = std::max(1, static_cast<int>(kChunk / (3 * static_cast<size_t>(W))));
    std::vector<Chunk> chunks;
    for (int y = 0; y < H; y += band) {
        Chunk c{};
        c.dst = reinterpret_cast<char*>(dst + (size_t)y * W);
        c.src = reinterpret_cast<const char*>(src + (size_t)y * srcStride);
        c.rows = std::min(band, H - y);
        c.W = W;
        c.srcStride = srcStride;
        chunks.push_back(c);
    }
    if ((size_t)W * H * 3 < kParallelMin) {
        for (const Chunk& c : chunks) bgr_rows(c);
        return;
    }
    Pool::get().run(chunks);
}

}  // namespace lumen
