// host_copy.h - parallel host memcpy (and BGR -> gray) for staging images and disparity maps
// through pinned buffers.
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

// OpenCV's fixed-point BGR -> gray (B 1868, G 9617, R 4899 in 14-bit fixed point, rounded) of a
// W x H interleaved image with row stride srcStride into a packed W x H gray image, in
// parallel bands of rows when the image is large enough to use the threads.
void parallel_bgr_to_gray(unsigned char* dst, const unsigned char* src, size_t srcStride, int W,
                          int H);

}  // namespace lumen
