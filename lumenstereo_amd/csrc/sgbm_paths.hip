// sgbm_paths.hip - per-pair SGM aggregation launch: all paths of a pair in one grid. The
// kernels of each slice K live in sgbm_paths_impl.h, instantiated by sgbm_paths_k*.hip.
#include "sgbm_device.h"

#include <algorithm>
#include <map>
#include <mutex>

namespace lumen {

int scanline_count(int path, int W, int H) { return scan_count(path, W, H); }

// Paths that run wider scanlines (32 lanes for 1-3, 64 for 4-6) under aggregation lane setting
// `wide` (Schedule::aggWide).
int wide_path_mask(int wide, int W, int H, int npaths) {
    switch (wide > 3 ? wide - 3 : wide) {
        case 1: return W >= H ? 0x3 : 0xC;  // the straight pair along the longer dimension
        case 2: return 0xF;                 // all four straight paths
        case 3: return (1 << npaths) - 1;   // every path
        default: return 0;
    }
}

// Dynamic LDS per aggregation block that lets at most n blocks share a CU: the LDS of a CU
// divided by n (at most 64 KiB, the most a launch may request without an attribute), which must
// still leave too little for an (n+1)-th block. Read once per device.
static hipError_t occupancy_lds(int n, size_t& lds) {
    static std::mutex mu;
    static std::map<int, int> perCu;  // device -> LDS bytes per CU
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    int cu = 0;
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = perCu.find(dev);
        if (it == perCu.end()) {
            e = hipDeviceGetAttribute(&cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev);
            if (e != hipSuccess) return e;
            perCu[dev] = cu;
        } else {
            cu = it->second;
        }
    }
    const size_t per = std::min<size_t>(65536, static_cast<size_t>(cu) / n) & ~size_t(255);
    if (per * (n + 1) <= static_cast<size_t>(cu)) return hipErrorInvalidValue;
    lds = per;
    return hipSuccess;
}

hipError_t launch_paths(const PairJob& j, uint16_t* trash, hipStream_t s, StageEvents ev) {
    PathsGrid g;
    const int w = j.sched.aggWide;
    if (w < 0 || w > 6 || (w >= 1 && w <= 3 && !wide_scanlines_ok(j.K, j.sched.aggPrefetch)) ||
        (w >= 4 && !xwide_scanlines_ok(j.K, j.sched.aggPrefetch)))
        return hipErrorInvalidValue;
    const int mask = wide_path_mask(w, j.W, j.H, j.npaths);
    g.wideMask = w <= 3 ? mask : 0;
    g.xwideMask = w >= 4 ? mask : 0;
    int waves = 0;
    for (int r = 0; r < j.npaths; ++r) {
        g.waveStart[r] = waves;
        const int per = (g.xwideMask >> r) & 1 ? 1 : (g.wideMask >> r) & 1 ? 2 : 4;
        waves += (scan_count(r, j.W, j.H) + per - 1) / per;
    }
    for (int r = j.npaths; r <= kMaxPaths; ++r) g.waveStart[r] = waves;
    const int blocks = (waves + 3) / 4;
    size_t lds = 0;
    if (j.sched.aggOccupancy != 0) {
        const hipError_t e = occupancy_lds(j.sched.aggOccupancy, lds);
        if (e != hipSuccess) return e;
    }
    switch (j.K) {
        case 4: return launch_paths_k<4>(j, g, blocks, lds, trash, s, ev);
        case 6: return launch_paths_k<6>(j, g, blocks, lds, trash, s, ev);
        case 8: return launch_paths_k<8>(j, g, blocks, lds, trash, s, ev);
        case 12: return launch_paths_k<12>(j, g, blocks, lds, trash, s, ev);
        case 16: return launch_paths_k<16>(j, g, blocks, lds, trash, s, ev);
        case 20: return launch_paths_k<20>(j, g, blocks, lds, trash, s, ev);
        case 24: return launch_paths_k<24>(j, g, blocks, lds, trash, s, ev);
        case 28: return launch_paths_k<28>(j, g, blocks, lds, trash, s, ev);
        case 32: return launch_paths_k<32>(j, g, blocks, lds, trash, s, ev);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace lumen
