/* lumen_device.h - device memory for callers of the C ABI: allocation, release, copies and
 * clearing of plain device buffers. It is what the header-only stereo::DeviceBuffer<T>
 * (include/stereo/device_buffer.h) is built on, so C++ callers keep the reference's RAII
 * buffer without including a GPU runtime header.
 *
 * Every call is synchronous with respect to the host, as the reference's cudaMalloc /
 * cudaMemcpy / cudaMemset calls are, and works on the current device of the calling thread.
 * Status codes are those of lumen_sgbm.h; lumen_last_error() gives the text of a failure.
 */
#ifndef LUMEN_DEVICE_H
#define LUMEN_DEVICE_H

#include <stddef.h>

#include "lumen_sgbm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Directions of lumen_device_copy. */
#define LUMEN_COPY_HOST_TO_DEVICE 0
#define LUMEN_COPY_DEVICE_TO_HOST 1
#define LUMEN_COPY_DEVICE_TO_DEVICE 2

/* Allocates `bytes` of device memory; *out receives the address, or NULL for 0 bytes.
 * LUMEN_ENOMEM when the device has no room. Replaces the cudaMalloc of
 * CudaBuffer::allocate (include/stereo/cuda_buffer.h:102-114). */
int lumen_device_alloc(size_t bytes, void** out);

/* Releases memory from lumen_device_alloc (NULL is a no-op). Replaces CudaBuffer::free
 * (include/stereo/cuda_buffer.h:119-125), which ignores the error of cudaFree; this call
 * reports it. */
int lumen_device_free(void* ptr);

/* Copies `bytes` in the given direction (LUMEN_COPY_*) and returns when the copy is done.
 * Replaces CudaBuffer::copyFrom / copyTo / copyFromDevice
 * (include/stereo/cuda_buffer.h:137-170). */
int lumen_device_copy(void* dst, const void* src, size_t bytes, int direction);

/* Sets `bytes` of device memory to zero and returns when done. Replaces CudaBuffer::zero
 * (include/stereo/cuda_buffer.h:175-179). */
int lumen_device_zero(void* ptr, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* LUMEN_DEVICE_H */
