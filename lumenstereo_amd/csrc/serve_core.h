// serve_core.h - host-side logic of the serving loop that does not touch the device: the
// device pool's allocator, admission by footprint, and latency percentiles.
//
// Pure C++ (no HIP), so it is exercised on the CPU with an injected backing allocator
// (tests/cpp/test_serve_core.cpp).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <deque>
#include <limits>
#include <map>
#include <unordered_map>
#include <utility>
#include <vector>

namespace lumen {

constexpr size_t kNoOffset = std::numeric_limits<size_t>::max();

// First-fit allocator of byte ranges inside [0, capacity): sizes are rounded up to `align`,
// freed ranges are merged with free neighbours.
class RegionAllocator {
public:
    explicit RegionAllocator(size_t capacity = 0, size_t align = 256) { reset(capacity, align); }

    void reset(size_t capacity, size_t align = 256) {
        align_ = align ? align : 1;
        capacity_ = capacity / align_ * align_;
        free_.clear();
        used_.clear();
        usedBytes_ = 0;
        if (capacity_) free_[0] = capacity_;
    }

    // Offset of a range of at least `bytes`, or kNoOffset when no free range is large enough.
    size_t alloc(size_t bytes) {
        const size_t need = round(std::max<size_t>(bytes, 1));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < need) continue;
            const size_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > need) free_[off + need] = len - need;
            used_[off] = need;
            usedBytes_ += need;
            return off;
        }
        return kNoOffset;
    }

    // Returns false (and changes nothing) for an offset that is not allocated.
    bool release(size_t off) {
        auto u = used_.find(off);
        if (u == used_.end()) return false;
        size_t start = off, len = u->second;
        used_.erase(u);
        usedBytes_ -= len;
        auto next = free_.lower_bound(start);
        if (next != free_.end() && start + len == next->first) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == start) {
                start = prev->first;
                len += prev->second;
                free_.erase(prev);
            }
        }
        free_[start] = len;
        return true;
    }

    size_t round(size_t bytes) const { return (bytes + align_ - 1) / align_ * align_; }
    size_t capacity() const { return capacity_; }
    size_t used_bytes() const { return usedBytes_; }
    size_t free_bytes() const { return capacity_ - usedBytes_; }
    size_t largest_free() const {
        size_t best = 0;
        for (const auto& f : free_) best = std::max(best, f.second);
        return best;
    }
    size_t free_ranges() const { return free_.size(); }
    size_t live_ranges() const { return used_.size(); }

private:
    size_t align_ = 256, capacity_ = 0, usedBytes_ = 0;
    std::map<size_t, size_t> free_;             // offset -> length, ordered by offset
    std::unordered_map<size_t, size_t> used_;   // offset -> length
};

// Where the pool's one big block comes from: hipMalloc on the device, anything in tests.
struct PoolBacking {
    virtual ~PoolBacking() = default;
    virtual void* acquire(size_t bytes) = 0;  // nullptr on failure
    virtual void release(void* p) = 0;
};

// One block from the backing, handed out in ranges. Never touches the memory itself.
class DevicePool {
public:
    DevicePool() = default;
    DevicePool(const DevicePool&) = delete;
    DevicePool& operator=(const DevicePool&) = delete;
    ~DevicePool() { close(); }

    bool open(PoolBacking* backing, size_t bytes) {
        close();
        base_ = backing ? static_cast<char*>(backing->acquire(bytes)) : nullptr;
        if (!base_) return false;
        backing_ = backing;
        ranges_.reset(bytes);
        highWater_ = 0;
        return true;
    }
    void close() {
        if (base_ && backing_) backing_->release(base_);
        base_ = nullptr;
        backing_ = nullptr;
        ranges_.reset(0);
    }

    void* alloc(size_t bytes) {
        const size_t off = ranges_.alloc(bytes);
        if (off == kNoOffset) return nullptr;
        highWater_ = std::max(highWater_, ranges_.used_bytes());
        return base_ + off;
    }
    bool release(void* p) {
        if (!base_ || !p) return false;
        const char* c = static_cast<const char*>(p);
        if (c < base_ || c >= base_ + ranges_.capacity()) return false;
        return ranges_.release(static_cast<size_t>(c - base_));
    }
    // Could a request of `bytes` ever be admitted (on an empty pool)?
    bool fits_ever(size_t bytes) const { return ranges_.round(bytes) <= ranges_.capacity(); }

    size_t capacity() const { return ranges_.capacity(); }
    size_t used_bytes() const { return ranges_.used_bytes(); }
    size_t high_water() const { return highWater_; }
    size_t live() const { return ranges_.live_ranges(); }
    const char* base() const { return base_; }

private:
    PoolBacking* backing_ = nullptr;
    char* base_ = nullptr;
    RegionAllocator ranges_;
    size_t highWater_ = 0;
};

// Admission by footprint. Requests wait in arrival order; admit() hands out pool ranges to the
// waiting requests that fit now. A request behind a head that does not fit may go first, but
// only `maxBypass` times for the same head: after that nothing is admitted until the head fits,
// so a large request is delayed by at most that many smaller ones.
class Admission {
public:
    struct Admitted {
        uint64_t id;
        void* region;
    };

    explicit Admission(DevicePool* pool = nullptr, int maxBypass = 8)
        : pool_(pool), maxBypass_(maxBypass) {}
    void attach(DevicePool* pool) { pool_ = pool; }

    void push(uint64_t id, size_t bytes) { queue_.push_back({id, bytes}); }

    std::vector<Admitted> admit() {
        std::vector<Admitted> out;
        if (!pool_) return out;
        auto it = queue_.begin();
        bool headBlocked = false;
        while (it != queue_.end()) {
            if (headBlocked && headBypassed_ >= maxBypass_) break;
            void* r = pool_->alloc(it->bytes);
            if (r) {
                out.push_back({it->id, r});
                if (headBlocked) ++headBypassed_;
                if (it == queue_.begin()) headBypassed_ = 0;
                it = queue_.erase(it);
            } else {
                if (it == queue_.begin()) headBlocked = true;
                ++it;
            }
        }
        return out;
    }

    // Removes a waiting request (e.g. when the server shuts down); false if it is not waiting.
    bool cancel(uint64_t id) {
        for (auto it = queue_.begin(); it != queue_.end(); ++it)
            if (it->id == id) {
                if (it == queue_.begin()) headBypassed_ = 0;
                queue_.erase(it);
                return true;
            }
        return false;
    }

    size_t waiting() const { return queue_.size(); }
    bool empty() const { return queue_.empty(); }
    int head_bypassed() const { return headBypassed_; }

private:
    struct Entry {
        uint64_t id;
        size_t bytes;
    };
    DevicePool* pool_;
    int maxBypass_;
    int headBypassed_ = 0;
    std::deque<Entry> queue_;
};

// Which ready requests (indices into `work`, in queue order) the next batch takes. A batch
// returns when its last pair finishes, so a small request batched with a large one waits for
// the large one: the batch starts with the oldest ready request and adds later ones only when
// their estimated work is within a factor `spread` of the first's either way, while the total
// stays within max(budget, first's work) and the count within maxBatch. Requests left out
// stay ready, in order, for the next free lane.
inline std::vector<size_t> pick_batch(const std::vector<double>& work, size_t maxBatch,
                                      double budget, double spread) {
    std::vector<size_t> out;
    if (work.empty() || maxBatch == 0) return out;
    const double first = work[0];
    const double cap = std::max(budget, first);
    double total = first;
    out.push_back(0);
    for (size_t i = 1; i < work.size() && out.size() < maxBatch; ++i) {
        const double w = work[i];
        if (w > first * spread || w * spread < first) continue;
        if (total + w > cap) continue;
        total += w;
        out.push_back(i);
    }
    return out;
}

// Nearest-rank percentile (q in [0, 100]) of a sample; 0 for an empty one.
inline double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    // The small slack keeps q/100*n from rounding up past an exact rank (99.9% of 1000 = 999).
    const double rank = std::ceil(q / 100.0 * static_cast<double>(v.size()) - 1e-9);
    const size_t idx = rank < 1.0 ? 0 : std::min(v.size() - 1, static_cast<size_t>(rank) - 1);
    return v[idx];
}

}  // namespace lumen
