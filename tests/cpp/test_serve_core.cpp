// Host-only tests of the serving loop's pool and admission logic (serve_core.h), with an
// injected backing allocator in place of hipMalloc, and of the staging copy pool
// (host_copy.cpp). Exit status 0 = all passed.
#include <cstdint>
#include <algorithm>
#include <cstdio>
#include <random>
#include <set>
#include <vector>

#include "host_copy.h"
#include "serve_core.h"

using namespace lumen;

static int g_failed = 0;
#define EXPECT(cond)                                                        \
    do {                                                                    \
        if (!(cond)) {                                                      \
            std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            ++g_failed;                                                     \
        }                                                                   \
    } while (0)

// Hands out fake addresses (never dereferenced) and records what was asked.
struct FakeBacking : PoolBacking {
    int acquired = 0, released = 0;
    size_t lastBytes = 0;
    bool failNext = false;
    void* acquire(size_t bytes) override {
        if (failNext) return nullptr;
        ++acquired;
        lastBytes = bytes;
        return reinterpret_cast<void*>(uintptr_t(0x7f0000000000ull));
    }
    void release(void*) override { ++released; }
};

static void allocator_basics() {
    RegionAllocator a(4096, 256);
    const size_t x = a.alloc(1);
    const size_t y = a.alloc(300);
    const size_t z = a.alloc(256);
    EXPECT(x == 0 && y == 256 && z == 768);
    EXPECT(a.used_bytes() == 1024 && a.free_bytes() == 3072);
    EXPECT(a.alloc(4096) == kNoOffset);
    EXPECT(a.release(y));
    EXPECT(!a.release(y));          // double free is refused
    EXPECT(!a.release(12345));      // unknown offset is refused
    EXPECT(a.alloc(512) == 256);    // first fit reuses the hole exactly
    EXPECT(a.release(x) && a.release(256) && a.release(z));
    EXPECT(a.free_ranges() == 1 && a.largest_free() == 4096 && a.used_bytes() == 0);
    RegionAllocator odd(1000, 256);  // capacity rounds down to the alignment
    EXPECT(odd.capacity() == 768 && odd.alloc(768) == 0 && odd.alloc(1) == kNoOffset);
}

// Random alloc/free against a byte map: ranges never overlap, stay in bounds, and the free
// list always describes exactly the bytes nobody holds.
static void allocator_model() {
    std::mt19937 rng(7);
    const size_t cap = 64 * 16;
    RegionAllocator a(cap, 16);
    std::vector<int> owner(cap, -1);
    std::vector<std::pair<size_t, size_t>> live;  // offset, rounded length
    for (int step = 0; step < 20000; ++step) {
        if (live.empty() || rng() % 2) {
            const size_t want = 1 + rng() % 200;
            const size_t off = a.alloc(want);
            const size_t len = (want + 15) / 16 * 16;
            // Brute force: is there any free run of len bytes at all?
            bool anyRun = false;
            for (size_t s = 0; s + len <= cap && !anyRun; s += 16) {
                bool ok = true;
                for (size_t b = s; b < s + len && ok; ++b) ok = owner[b] < 0;
                anyRun = ok;
            }
            if (off == kNoOffset) {
                EXPECT(!anyRun);
                continue;
            }
            EXPECT(anyRun);
            EXPECT(off % 16 == 0 && off + len <= cap);
            for (size_t b = off; b < off + len; ++b) {
                EXPECT(owner[b] < 0);
                owner[b] = step;
            }
            live.push_back({off, len});
        } else {
            const size_t i = rng() % live.size();
            EXPECT(a.release(live[i].first));
            for (size_t b = live[i].first; b < live[i].first + live[i].second; ++b) owner[b] = -1;
            live.erase(live.begin() + i);
        }
        size_t held = 0;
        for (auto& l : live) held += l.second;
        EXPECT(a.used_bytes() == held);
        if (g_failed) return;
    }
    for (auto& l : live) a.release(l.first);
    EXPECT(a.free_ranges() == 1 && a.free_bytes() == cap);
}

static void pool_with_injected_backing() {
    FakeBacking fb;
    DevicePool pool;
    fb.failNext = true;
    EXPECT(!pool.open(&fb, 1 << 20));
    fb.failNext = false;
    EXPECT(pool.open(&fb, 1 << 20));
    EXPECT(fb.acquired == 1 && fb.lastBytes == (1u << 20));
    char* base = const_cast<char*>(pool.base());
    void* p = pool.alloc(1000);
    void* q = pool.alloc(5000);
    EXPECT(p == base && q == base + 1024);
    EXPECT(pool.used_bytes() == 1024 + 5120 && pool.live() == 2);
    EXPECT(!pool.release(base + 8));       // not the start of a range
    EXPECT(!pool.release(base + (2 << 20)));  // outside the pool
    EXPECT(pool.release(p) && pool.release(q));
    EXPECT(pool.high_water() == 1024 + 5120 && pool.used_bytes() == 0);
    EXPECT(pool.fits_ever(1 << 20) && !pool.fits_ever((1 << 20) + 1));
    EXPECT(pool.alloc((1 << 20) + 1) == nullptr);
    pool.close();
    EXPECT(fb.released == 1);
}

static void admission_fifo_and_bypass() {
    FakeBacking fb;
    DevicePool pool;
    pool.open(&fb, 10 * 256);
    Admission adm(&pool, 2);
    // Pool holds 10 units. Request 1 takes 8; request 2 (6 units) is then blocked; 3..6 (1 unit
    // each) may pass it, but only twice.
    adm.push(1, 8 * 256);
    auto a = adm.admit();
    EXPECT(a.size() == 1 && a[0].id == 1);
    void* big = a[0].region;
    for (uint64_t id = 2; id <= 6; ++id) adm.push(id, id == 2 ? 6 * 256 : 256);
    a = adm.admit();
    EXPECT(a.size() == 2 && a[0].id == 3 && a[1].id == 4);
    EXPECT(adm.head_bypassed() == 2 && adm.waiting() == 3);
    std::vector<Admission::Admitted> small = a;
    EXPECT(adm.admit().empty());  // the head still does not fit; no more bypass
    pool.release(big);
    a = adm.admit();
    // Head (6 units) fits now beside 3 and 4, then 5 and 6: 2 + 6 + 1 + 1 = 10 units.
    EXPECT(a.size() == 3 && a[0].id == 2 && a[1].id == 5 && a[2].id == 6);
    EXPECT(adm.head_bypassed() == 0 && adm.empty());
    for (auto& x : a) pool.release(x.region);
    for (auto& x : small) pool.release(x.region);
    EXPECT(pool.used_bytes() == 0);

    // maxBypass 0: strict arrival order.
    Admission fifo(&pool, 0);
    fifo.push(10, 10 * 256);
    fifo.push(11, 256);
    void* hold = pool.alloc(256);
    EXPECT(fifo.admit().empty());
    pool.release(hold);
    a = fifo.admit();
    EXPECT(a.size() == 1 && a[0].id == 10);
    for (auto& x : a) pool.release(x.region);
    a = fifo.admit();
    EXPECT(a.size() == 1 && a[0].id == 11);
    for (auto& x : a) pool.release(x.region);
    EXPECT(fifo.empty());

    Admission c(&pool, 4);
    c.push(20, 256);
    c.push(21, 256);
    EXPECT(c.cancel(20) && !c.cancel(20) && c.waiting() == 1);
}

// Discrete-event run of admission + completion with random footprints and durations: the pool
// never holds more than its capacity, every request is admitted exactly once, and the request
// at the head of the queue is passed at most maxBypass times.
static void admission_simulation() {
    std::mt19937 rng(11);
    FakeBacking fb;
    DevicePool pool;
    const size_t cap = 1000 * 256;
    pool.open(&fb, cap);
    const int maxBypass = 3;
    Admission adm(&pool, maxBypass);
    struct Running {
        uint64_t id;
        void* region;
        int end;
    };
    std::vector<Running> running;
    std::set<uint64_t> admitted;
    uint64_t next = 1;
    const uint64_t total = 3000;
    for (int t = 0; admitted.size() < total && t < 200000; ++t) {
        if (next <= total && rng() % 3 == 0) {
            const size_t units = rng() % 10 == 0 ? 400 + rng() % 500 : 1 + rng() % 60;
            adm.push(next++, units * 256);
        }
        for (size_t i = 0; i < running.size();) {
            if (running[i].end <= t) {
                EXPECT(pool.release(running[i].region));
                running.erase(running.begin() + i);
            } else {
                ++i;
            }
        }
        for (auto& a : adm.admit()) {
            EXPECT(admitted.insert(a.id).second);
            running.push_back({a.id, a.region, t + 1 + static_cast<int>(rng() % 20)});
        }
        EXPECT(pool.used_bytes() <= cap);
        EXPECT(adm.head_bypassed() <= maxBypass);
        if (g_failed) return;
    }
    EXPECT(admitted.size() == total);
    EXPECT(pool.high_water() <= cap);
}

static void percentiles() {
    std::vector<double> v;
    for (int i = 1; i <= 1000; ++i) v.push_back(i);
    EXPECT(percentile(v, 50) == 500 && percentile(v, 99) == 990 && percentile(v, 99.9) == 999);
    EXPECT(percentile(v, 100) == 1000 && percentile(v, 0) == 1);
    EXPECT(percentile({}, 50) == 0);
    EXPECT(percentile({3, 1, 2}, 50) == 2);
}

// The staging copy pool: many jobs of odd sizes, large enough to use the threads, twice in a
// row (the pool is reused), and strided rows.
static void parallel_copies() {
    std::mt19937 rng(3);
    std::vector<std::vector<uint8_t>> src, dst;
    std::vector<CopyJob> jobs;
    for (int i = 0; i < 40; ++i) {
        const size_t n = 1 + rng() % (700 * 1024);
        src.emplace_back(n);
        for (auto& b : src.back()) b = static_cast<uint8_t>(rng());
        dst.emplace_back(n, 0);
    }
    for (int rep = 0; rep < 2; ++rep) {
        jobs.clear();
        for (size_t i = 0; i < src.size(); ++i) {
            std::fill(dst[i].begin(), dst[i].end(), 0);
            jobs.push_back({dst[i].data(), src[i].data(), src[i].size()});
        }
        parallel_copy(jobs);
        for (size_t i = 0; i < src.size(); ++i) EXPECT(dst[i] == src[i]);
    }
    const int W = 1000, H = 3000, sp = 1200, dp = 1100;
    std::vector<uint8_t> a((size_t)sp * H), b((size_t)dp * H, 7);
    for (auto& v : a) v = static_cast<uint8_t>(rng());
    parallel_copy_rows(b.data(), dp, a.data(), sp, W, H);
    bool ok = true;
    for (int y = 0; y < H && ok; ++y)
        for (int x = 0; x < dp; ++x) {
            const uint8_t want = x < W ? a[(size_t)y * sp + x] : 7;
            if (b[(size_t)y * dp + x] != want) ok = false;
        }
    EXPECT(ok);
}

// BGR -> gray in bands on the pool, against the formula pixel by pixel: below the parallel
// threshold, above it, with a padded source stride, twice in a row.
static void parallel_bgr() {
    std::mt19937 rng(5);
    const int sizes[][3] = {{7, 5, 21}, {640, 480, 1920}, {1921, 1081, 5800}, {1000, 3, 3000}};
    for (int rep = 0; rep < 2; ++rep)
        for (const auto& sz : sizes) {
            const int W = sz[0], H = sz[1];
            const size_t stride = static_cast<size_t>(sz[2]);
            std::vector<uint8_t> src(stride * H);
            for (auto& v : src) v = static_cast<uint8_t>(rng());
            std::vector<uint8_t> dst((size_t)W * H, 0);
            parallel_bgr_to_gray(dst.data(), src.data(), stride, W, H);
            bool ok = true;
            for (int y = 0; y < H && ok; ++y)
                for (int x = 0; x < W; ++x) {
                    const uint8_t* p = src.data() + (size_t)y * stride + 3 * x;
                    const int want = (p[0] * 1868 + p[1] * 9617 + p[2] * 4899 + 8192) >> 14;
                    if (dst[(size_t)y * W + x] != want) ok = false;
                }
            EXPECT(ok);
        }
}

static void batch_picking() {
    // Oldest first; only requests within 4x of the first's work; total within the budget.
    std::vector<double> w = {10, 100, 12, 3, 30, 41, 9, 11};
    auto p = pick_batch(w, 64, 50, 4.0);
    EXPECT((p == std::vector<size_t>{0, 2, 3, 6, 7}));  // 10+12+3+9+11 = 45; 30 would exceed 50
    p = pick_batch(w, 2, 1000, 4.0);
    EXPECT((p == std::vector<size_t>{0, 2}));         // count cap
    w = {100, 10, 30, 400, 90};
    p = pick_batch(w, 64, 50, 4.0);
    EXPECT((p == std::vector<size_t>{0}));            // a large first request: 90 would exceed max(50, 100)
    p = pick_batch(w, 64, 500, 4.0);
    EXPECT((p == std::vector<size_t>{0, 2, 4}));      // 10 too small; 400 would exceed 500
    EXPECT(pick_batch({}, 64, 50, 4.0).empty());
    EXPECT((pick_batch({5}, 64, 1, 4.0) == std::vector<size_t>{0}));  // always at least one
}

int main() {
    struct {
        const char* name;
        void (*fn)();
    } cases[] = {
        {"allocator_basics", allocator_basics},
        {"allocator_model", allocator_model},
        {"pool_with_injected_backing", pool_with_injected_backing},
        {"admission_fifo_and_bypass", admission_fifo_and_bypass},
        {"admission_simulation", admission_simulation},
        {"percentiles", percentiles},
        {"parallel_copies", parallel_copies},
        {"parallel_bgr", parallel_bgr},
        {"batch_picking", batch_picking},
    };
    for (auto& c : cases) {
        const int before = g_failed;
        c.fn();
        std::printf("%s %s\n", g_failed == before ? "ok  " : "FAIL", c.name);
    }
    return g_failed ? 1 : 0;
}
