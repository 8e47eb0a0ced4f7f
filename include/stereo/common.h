#pragma once
// Error types and helpers of include/stereo/common.h of the reference (:31-145).
// CudaException becomes GpuException: it carries the library status code instead of a
// cudaError_t. StereoException and CalibrationException keep their names and messages.
#include <sstream>
#include <stdexcept>
#include <string>

namespace stereo {

class GpuException : public std::runtime_error {
public:
    GpuException(int code, const std::string& what)
        : std::runtime_error(format(code, what)), error_code(code) {}
    int error_code;

private:
    static std::string format(int code, const std::string& what) {
        std::ostringstream ss;
        ss << "GPU error " << code << ": " << what;
        return ss.str();
    }
};

class StereoException : public std::runtime_error {
public:
    explicit StereoException(const std::string& msg) : std::runtime_error(msg) {}
};

class CalibrationException : public StereoException {
public:
    explicit CalibrationException(const std::string& msg)
        : StereoException("Calibration error: " + msg) {}
};

inline int divUp(int a, int b) { return (a + b - 1) / b; }

}  // namespace stereo
