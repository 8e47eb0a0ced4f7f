// sgbm_paths.hip - per-pair SGM aggregation launch: all paths of a pair in one grid.
//
// One instantiation per (slice K, padding, prefetch depth): the prefetch depth is a schedule
// choice (Schedule::aggPrefetch), so each pair runs the instantiation its schedule names.
#include "sgbm_device.h"

namespace lumen {

int scanline_count(int path, int W, int H) { return scan_count(path, W, H); }

// Wave w takes the next scanlines of the path whose range of waves contains w (paths in order,
// so the long horizontal scanlines start first): four of 16 lanes each, or, for the paths in
// wideMask, two of 32 lanes each (half the disparities per lane, so about a third fewer
// instructions per step on each scanline's dependency chain, for twice the waves).
struct PathsGrid {
    int waveStart[kMaxPaths + 1];
    int wideMask;  // bit r: path r runs 32-lane scanlines
};

template <int K, bool PAD, int PF>
__global__ __launch_bounds__(256) void sgm_paths_kernel(PairJob j, PathsGrid g, uint16_t* trash) {
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wave >= g.waveStart[j.npaths]) return;
    int path = 0;
#pragma unroll
    for (int r = 1; r < kMaxPaths; ++r)
        if (r < j.npaths && wave >= g.waveStart[r]) path = r;
    const bool wide = (g.wideMask >> path) & 1;
    const int s0 = (wave - g.waveStart[path]) * (wide ? 2 : 4);
    // The launch ends when its longest scanlines end. Priority 1: the longer straight path
    // first, then the other straight path, diagonals last; priority 2: by scanline length.
    if (j.sched.aggPriority == 1) {
        const bool longest = j.W >= j.H ? path < 2 : (path >= 2 && path < 4);
        if (longest)
            __builtin_amdgcn_s_setprio(3);
        else if (path < 4)
            __builtin_amdgcn_s_setprio(2);
        else
            __builtin_amdgcn_s_setprio(1);
    } else if (j.sched.aggPriority == 2) {
        set_priority_for_length(scan_length(path, s0, j.W, j.H));
    }
    if constexpr (wide_scanlines_ok(K, PF)) {
        if (wide) {
            paths_wave<K / 2, PAD, PF, 32>(j, path, s0, trash);
            return;
        }
    }
    paths_wave<K, PAD, PF>(j, path, s0, trash);
}

template <int K, int PF>
static hipError_t launch_kp(const PairJob& j, const PathsGrid& g, int blocks, uint16_t* trash,
                            hipStream_t s, StageEvents ev) {
    if constexpr (!valid_prefetch(K, PF)) {
        return hipErrorInvalidValue;
    } else {
        if (j.D == j.Dp)
            hipExtLaunchKernelGGL((sgm_paths_kernel<K, false, PF>), dim3(blocks), dim3(256), 0, s,
                                  ev.start, ev.stop, 0, j, g, trash);
        else
            hipExtLaunchKernelGGL((sgm_paths_kernel<K, true, PF>), dim3(blocks), dim3(256), 0, s,
                                  ev.start, ev.stop, 0, j, g, trash);
        return hipGetLastError();
    }
}

template <int K>
static hipError_t launch_k(const PairJob& j, const PathsGrid& g, int blocks, uint16_t* trash,
                           hipStream_t s, StageEvents ev) {
    switch (j.sched.aggPrefetch) {
        case 4: return launch_kp<K, 4>(j, g, blocks, trash, s, ev);
        case 8: return launch_kp<K, 8>(j, g, blocks, trash, s, ev);
        case 16: return launch_kp<K, 16>(j, g, blocks, trash, s, ev);
        default: return hipErrorInvalidValue;
    }
}

// Paths that run 32-lane scanlines under aggregation lane setting `wide` (Schedule::aggWide).
int wide_path_mask(int wide, int W, int H, int npaths) {
    switch (wide) {
        case 1: return W >= H ? 0x3 : 0xC;  // the straight pair along the longer dimension
        case 2: return 0xF;                 // all four straight paths
        case 3: return (1 << npaths) - 1;   // every path
        default: return 0;
    }
}

hipError_t launch_paths(const PairJob& j, uint16_t* trash, hipStream_t s, StageEvents ev) {
    PathsGrid g;
    if (j.sched.aggWide != 0 && !wide_scanlines_ok(j.K, j.sched.aggPrefetch))
        return hipErrorInvalidValue;
    g.wideMask = wide_path_mask(j.sched.aggWide, j.W, j.H, j.npaths);
    int waves = 0;
    for (int r = 0; r < j.npaths; ++r) {
        g.waveStart[r] = waves;
        const int per = (g.wideMask >> r) & 1 ? 2 : 4;
        waves += (scan_count(r, j.W, j.H) + per - 1) / per;
    }
    for (int r = j.npaths; r <= kMaxPaths; ++r) g.waveStart[r] = waves;
    const int blocks = (waves + 3) / 4;
    switch (j.K) {
        case 4: return launch_k<4>(j, g, blocks, trash, s, ev);
        case 8: return launch_k<8>(j, g, blocks, trash, s, ev);
        case 12: return launch_k<12>(j, g, blocks, trash, s, ev);
        case 16: return launch_k<16>(j, g, blocks, trash, s, ev);
        case 20: return launch_k<20>(j, g, blocks, trash, s, ev);
        case 24: return launch_k<24>(j, g, blocks, trash, s, ev);
        case 28: return launch_k<28>(j, g, blocks, trash, s, ev);
        case 32: return launch_k<32>(j, g, blocks, trash, s, ev);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace lumen
