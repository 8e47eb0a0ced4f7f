// device_mem.hip - the C ABI of include/lumen_device.h: plain device allocations, copies and
// clearing for callers that hold their own buffers (stereo::DeviceBuffer<T>). Each call waits
// for its own work, as the blocking runtime calls the reference's CudaBuffer makes.
#include "lumen_device.h"

#include <hip/hip_runtime.h>

#include <string>

namespace lumen {
void set_last_error(const std::string& msg);
}

namespace {

int fail(hipError_t e, const char* what) {
    lumen::set_last_error(std::string(what) + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? LUMEN_ENOMEM : LUMEN_EDEVICE;
}

}  // namespace

extern "C" int lumen_device_alloc(size_t bytes, void** out) {
    if (!out) return (lumen::set_last_error("out is NULL"), LUMEN_EINVAL);
    *out = nullptr;
    if (bytes == 0) return LUMEN_OK;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // an allocation failure must not stick to the next call
        return fail(e, "hipMalloc");
    }
    *out = p;
    return LUMEN_OK;
}

extern "C" int lumen_device_free(void* ptr) {
    if (!ptr) return LUMEN_OK;
    hipError_t e = hipFree(ptr);
    return e == hipSuccess ? LUMEN_OK : fail(e, "hipFree");
}

extern "C" int lumen_device_copy(void* dst, const void* src, size_t bytes, int direction) {
    if (bytes == 0) return LUMEN_OK;
    if (!dst || !src) return (lumen::set_last_error("NULL buffer"), LUMEN_EINVAL);
    hipMemcpyKind kind;
    switch (direction) {
        case LUMEN_COPY_HOST_TO_DEVICE: kind = hipMemcpyHostToDevice; break;
        case LUMEN_COPY_DEVICE_TO_HOST: kind = hipMemcpyDeviceToHost; break;
        case LUMEN_COPY_DEVICE_TO_DEVICE: kind = hipMemcpyDeviceToDevice; break;
        default: return (lumen::set_last_error("unknown copy direction"), LUMEN_EINVAL);
    }
    hipError_t e = hipMemcpy(dst, src, bytes, kind);
    // A device-to-device copy may return before it is done; the call promises completion.
    if (e == hipSuccess && kind == hipMemcpyDeviceToDevice) e = hipStreamSynchronize(nullptr);
    return e == hipSuccess ? LUMEN_OK : fail(e, "hipMemcpy");
}

extern "C" int lumen_device_zero(void* ptr, size_t bytes) {
    if (bytes == 0) return LUMEN_OK;
    if (!ptr) return (lumen::set_last_error("NULL buffer"), LUMEN_EINVAL);
    hipError_t e = hipMemset(ptr, 0, bytes);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    return e == hipSuccess ? LUMEN_OK : fail(e, "hipMemset");
}
