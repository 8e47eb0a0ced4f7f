// host_copy.h - parallel host memcpy for staging images and disparity maps through pinned
// buffers.
//
// One thread copies about 10 GB/s; a batch moves a few hundred MB between the caller's
// buffers and the pinned staging, which single-threaded costs as much as the batch's compute.
// Copies are split into 1 MiB chunks taken by a small persistent pool plus the caller.
#pragma once
#include <cstddef>
#include <vector>

namespace lumen {

struct CopyJob {
    void* dst;
    const void* src;
    size_t bytes;
};

// Runs every job (any sizes, non-overlapping), in parallel when there is enough to copy.
// Thread-safe: concurrent callers share the pool.
void parallel_copy(const std::vector<CopyJob>& jobs);

// Row-by-row copy of an image (strided source/destination) as one parallel operation.
void parallel_copy_rows(void* dst, size_t dstStride, const void* src, size_t srcStride,
                        size_t rowBytes, int rows);

}  // namespace lumen
