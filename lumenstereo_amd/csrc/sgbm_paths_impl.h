// sgbm_paths_impl.h - the per-pair SGM aggregation kernel and its launcher for one slice K.
//
// One instantiation per (slice K, padding, prefetch depth): the prefetch depth is a schedule
// choice (Schedule::aggPrefetch), so each pair runs the instantiation its schedule names. The
// slices are compiled in separate translation units (sgbm_paths_k*.hip) so that they build in
// parallel; sgbm_paths.hip picks the slice.
#pragma once
#include "sgbm_device.h"

namespace lumen {

// Wave w takes the next scanlines of the path whose range of waves contains w (paths in order,
// so the long horizontal scanlines start first): four of 16 lanes each, or, for the paths in
// wideMask, two of 32 lanes each (half the disparities per lane, so about a third fewer
// instructions per step on each scanline's dependency chain, for twice the waves), or for those
// in xwideMask one of 64 lanes (a quarter of the disparities per lane).
template <int K, bool PAD, int PF>
__global__ __launch_bounds__(256) void sgm_paths_kernel(PairJob j, PathsGrid g, uint16_t* trash) {
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wave >= g.waveStart[j.npaths]) return;
    int path = 0;
#pragma unroll
    for (int r = 1; r < kMaxPaths; ++r)
        if (r < j.npaths && wave >= g.waveStart[r]) path = r;
    const bool wide = (g.wideMask >> path) & 1, xwide = (g.xwideMask >> path) & 1;
    const int s0 = (wave - g.waveStart[path]) * (xwide ? 1 : wide ? 2 : 4);
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
    // The horizontal paths run with their step known at compile time (immediate offsets).
    if constexpr (xwide_scanlines_ok(K, PF)) {
        if (xwide) {
            if (path == 0)
                paths_wave<K / 4, PAD, PF, 64, 1>(j, path, s0, trash);
            else if (path == 1)
                paths_wave<K / 4, PAD, PF, 64, -1>(j, path, s0, trash);
            else
                paths_wave<K / 4, PAD, PF, 64>(j, path, s0, trash);
            return;
        }
    }
    if constexpr (wide_scanlines_ok(K, PF)) {
        if (wide) {
            if (path == 0)
                paths_wave<K / 2, PAD, PF, 32, 1>(j, path, s0, trash);
            else if (path == 1)
                paths_wave<K / 2, PAD, PF, 32, -1>(j, path, s0, trash);
            else
                paths_wave<K / 2, PAD, PF, 32>(j, path, s0, trash);
            return;
        }
    }
    if (path == 0)
        paths_wave<K, PAD, PF, 16, 1>(j, path, s0, trash);
    else if (path == 1)
        paths_wave<K, PAD, PF, 16, -1>(j, path, s0, trash);
    else
        paths_wave<K, PAD, PF>(j, path, s0, trash);
}

template <int K, int PF>
static hipError_t launch_kp(const PairJob& j, const PathsGrid& g, int blocks, size_t lds,
                            uint16_t* trash, hipStream_t s, StageEvents ev) {
    if constexpr (!valid_prefetch(K, PF)) {
        return hipErrorInvalidValue;
    } else {
        if (j.D == j.Dp)
            hipExtLaunchKernelGGL((sgm_paths_kernel<K, false, PF>), dim3(blocks), dim3(256), lds, s,
                                  ev.start, ev.stop, 0, j, g, trash);
        else
            hipExtLaunchKernelGGL((sgm_paths_kernel<K, true, PF>), dim3(blocks), dim3(256), lds, s,
                                  ev.start, ev.stop, 0, j, g, trash);
        return hipGetLastError();
    }
}

// lds: dynamic LDS bytes per block, reserved only to cap the blocks per CU (the kernel uses none).
template <int K>
hipError_t launch_paths_k(const PairJob& j, const PathsGrid& g, int blocks, size_t lds,
                          uint16_t* trash, hipStream_t s, StageEvents ev) {
    switch (j.sched.aggPrefetch) {
        case 4: return launch_kp<K, 4>(j, g, blocks, lds, trash, s, ev);
        case 8: return launch_kp<K, 8>(j, g, blocks, lds, trash, s, ev);
        case 16: return launch_kp<K, 16>(j, g, blocks, lds, trash, s, ev);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace lumen
