#pragma once
// RAII device buffer: the counterpart of the reference's CudaBuffer<T>
// (include/stereo/cuda_buffer.h:49-218), header-only over the C ABI of lumen_device.h, so a
// caller needs no GPU runtime header. Same members and semantics: move-only, sizes in
// elements, allocate() keeps an allocation of the same size, resize() reallocates (contents
// are not kept), copyFrom/copyTo default to the whole buffer and refuse counts past its end
// with StereoException, and runtime failures throw GpuException (the reference's
// CudaException). The reference name is not kept, as for the exception: callers switch
// CudaBuffer to DeviceBuffer.
#include <cstddef>

#include "common.h"
#include "lumen_device.h"

namespace stereo {

namespace detail {
inline void device_check(int rc, const char* what) {
    if (rc != LUMEN_OK) {
        const char* msg = lumen_last_error();
        throw GpuException(rc, std::string(what) + ": " + (msg ? msg : ""));
    }
}
}  // namespace detail

template <typename T>
class DeviceBuffer {
public:
    DeviceBuffer() : data_(nullptr), size_(0) {}
    explicit DeviceBuffer(size_t size) : data_(nullptr), size_(0) { allocate(size); }
    ~DeviceBuffer() { free(); }

    DeviceBuffer(DeviceBuffer&& other) noexcept : data_(other.data_), size_(other.size_) {
        other.data_ = nullptr;
        other.size_ = 0;
    }
    DeviceBuffer& operator=(DeviceBuffer&& other) noexcept {
        if (this != &other) {
            free();
            data_ = other.data_;
            size_ = other.size_;
            other.data_ = nullptr;
            other.size_ = 0;
        }
        return *this;
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;

    // Allocates `size` elements; a buffer already of that size is kept as it is.
    void allocate(size_t size) {
        if (size == size_ && data_ != nullptr) return;
        free();
        if (size > 0) {
            void* p = nullptr;
            detail::device_check(lumen_device_alloc(size * sizeof(T), &p), "DeviceBuffer::allocate");
            data_ = static_cast<T*>(p);
            size_ = size;
        }
    }

    // Releases the memory. Like the reference, errors of the release are not raised (this
    // runs in the destructor).
    void free() {
        if (data_ != nullptr) {
            (void)lumen_device_free(data_);
            data_ = nullptr;
            size_ = 0;
        }
    }

    void resize(size_t newSize) {
        if (newSize != size_) allocate(newSize);
    }

    // Host -> device; count 0 means the whole buffer.
    void copyFrom(const T* hostPtr, size_t count = 0) {
        if (count == 0) count = size_;
        if (count > size_) throw StereoException("DeviceBuffer::copyFrom: count exceeds buffer size");
        detail::device_check(lumen_device_copy(data_, hostPtr, count * sizeof(T),
                                               LUMEN_COPY_HOST_TO_DEVICE),
                             "DeviceBuffer::copyFrom");
    }

    // Device -> host; count 0 means the whole buffer.
    void copyTo(T* hostPtr, size_t count = 0) const {
        if (count == 0) count = size_;
        if (count > size_) throw StereoException("DeviceBuffer::copyTo: count exceeds buffer size");
        detail::device_check(lumen_device_copy(hostPtr, data_, count * sizeof(T),
                                               LUMEN_COPY_DEVICE_TO_HOST),
                             "DeviceBuffer::copyTo");
    }

    // Device -> device: the whole of `other` into the start of this buffer.
    void copyFromDevice(const DeviceBuffer<T>& other) {
        if (other.size() > size_)
            throw StereoException("DeviceBuffer::copyFromDevice: source larger than destination");
        detail::device_check(lumen_device_copy(data_, other.data(), other.size() * sizeof(T),
                                               LUMEN_COPY_DEVICE_TO_DEVICE),
                             "DeviceBuffer::copyFromDevice");
    }

    void zero() {
        if (data_ != nullptr)
            detail::device_check(lumen_device_zero(data_, size_ * sizeof(T)), "DeviceBuffer::zero");
    }

    T* data() { return data_; }
    const T* data() const { return data_; }
    size_t size() const { return size_; }
    size_t bytes() const { return size_ * sizeof(T); }
    bool empty() const { return data_ == nullptr || size_ == 0; }
    operator T*() { return data_; }
    operator const T*() const { return data_; }

private:
    T* data_;
    size_t size_;
};

// An image-sized buffer of width * height * channels elements (include/stereo/cuda_buffer.h:205-216).
template <typename T>
DeviceBuffer<T> make2DBuffer(int width, int height, int channels = 1) {
    return DeviceBuffer<T>(static_cast<size_t>(width) * height * channels);
}

}  // namespace stereo
