// scene_common.h - host helpers of the stages either side of the matcher (rectification,
// depth, colour map): a process-wide device staging area for the host-buffer entry points.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>

#include "lumen_sgbm.h"

namespace lumen {

void set_last_error(const std::string& msg);

inline int scene_hip_fail(hipError_t e, const char* what) {
    set_last_error(std::string(what) + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? LUMEN_ENOMEM : LUMEN_EDEVICE;
}

// Grow-only device buffers and one stream, shared by every host-buffer call of this process on
// one device and serialised by `mu`. The reference allocates its CudaBuffers on every call
// (src/rectification.cu:383-395, src/depth_map.cu:182); here they are kept. Never freed: they
// live until the process ends.
struct SceneStaging {
    std::mutex mu;
    int device = -1;
    hipStream_t stream = nullptr;
    void* buf[4] = {};
    size_t cap[4] = {};

    // Locks must be held. Makes buffer i at least `bytes` long on the current device.
    hipError_t ensure(int i, size_t bytes) {
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) return e;
        if (dev != device) {
            // Another device: start over there (the old device's buffers stay allocated).
            for (int k = 0; k < 4; ++k) {
                buf[k] = nullptr;
                cap[k] = 0;
            }
            stream = nullptr;
            device = dev;
        }
        if (!stream) {
            e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
            if (e != hipSuccess) {
                stream = nullptr;
                return e;
            }
        }
        if (bytes <= cap[i] && buf[i]) return hipSuccess;
        if (buf[i]) {
            e = hipFree(buf[i]);
            buf[i] = nullptr;
            cap[i] = 0;
            if (e != hipSuccess) return e;
        }
        e = hipMalloc(&buf[i], bytes ? bytes : 1);
        if (e != hipSuccess) {
            buf[i] = nullptr;
            return e;
        }
        cap[i] = bytes;
        return hipSuccess;
    }
};

SceneStaging& scene_staging();

}  // namespace lumen
