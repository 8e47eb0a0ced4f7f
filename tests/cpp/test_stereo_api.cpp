// C++ API tests mirroring the reference's tests/test_sgbm.cpp (StereoParams, validation and
// exception cases, output shape/type on GPU, depth) and tests/test_rectification.cpp
// (CameraParams, parameter files, Middlebury calibration, rectification). gtest is not
// available, so this is a minimal
// runner: `test_stereo_api` runs the host-only cases, `test_stereo_api --gpu` adds the cases
// that compute on the GPU. Exit status 0 = all passed.
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <utility>
#include <vector>

#include <cmath>
#include <fstream>

#include "stereo/depth_map.h"
#include "stereo/device_buffer.h"
#include "stereo/rectification.h"
#include "stereo/sgbm_gpu.h"

using namespace stereo;

static int g_failed = 0;
#define EXPECT(cond)                                                             \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);      \
            ++g_failed;                                                          \
        }                                                                        \
    } while (0)

template <typename E, typename F>
static bool throws(F&& f) {
    try {
        f();
    } catch (const E&) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}

// --- tests/test_sgbm.cpp:28-75 ---------------------------------------------------------
static void DefaultValues() {
    StereoParams p;
    EXPECT(p.minDisparity == 0);
    EXPECT(p.maxDisparity == 128);
    EXPECT(p.blockSize == 5);
    EXPECT(p.isValid());
}
static void InvalidBlockSize() {
    StereoParams p;
    p.blockSize = 4;
    EXPECT(!p.isValid());
    p.blockSize = 0;
    EXPECT(!p.isValid());
    p.blockSize = -1;
    EXPECT(!p.isValid());
}
static void InvalidDisparity() {
    StereoParams p;
    p.minDisparity = 100;
    p.maxDisparity = 50;
    EXPECT(!p.isValid());
    p.minDisparity = -10;
    p.maxDisparity = 128;
    EXPECT(!p.isValid());
}
static void InvalidPenalties() {
    StereoParams p;
    p.P1 = 100;
    p.P2 = 50;
    EXPECT(!p.isValid());
}
static void AutoPenalties() {
    StereoParams p;
    p.blockSize = 5;
    p.autoComputePenalties(1);
    EXPECT(p.P1 == 200);
    EXPECT(p.P2 == 800);
    EXPECT(p.P2 > p.P1);
}
static void DefaultsMatchCAbi() {
    lumen_stereo_params c;
    EXPECT(lumen_sgbm_default_params(&c) == LUMEN_OK);
    const lumen_stereo_params d = StereoParams().toC();
    EXPECT(c.minDisparity == d.minDisparity && c.maxDisparity == d.maxDisparity &&
           c.blockSize == d.blockSize && c.P1 == d.P1 && c.P2 == d.P2 &&
           c.preFilterCap == d.preFilterCap && c.uniquenessRatio == d.uniquenessRatio &&
           c.disp12MaxDiff == d.disp12MaxDiff && c.speckleWindowSize == d.speckleWindowSize &&
           c.speckleRange == d.speckleRange && c.numDirections == d.numDirections &&
           c.matchingCostMode == d.matchingCostMode);
}

// --- tests/test_sgbm.cpp:359-406 (thrown before any device work) ------------------------
static void EmptyInput() {
    StereoSGBM sgbm;
    std::vector<int16_t> out(1);
    EXPECT(throws<StereoException>([&] { sgbm.compute(ImageView{}, ImageView{}, out.data()); }));
    EXPECT(throws<StereoException>([&] { sgbm.compute(nullptr, nullptr, 0, 0, out.data()); }));
}
static void SizeMismatch() {
    std::vector<uint8_t> a(100 * 100), b(100 * 200);
    std::vector<int16_t> out(100 * 200);
    StereoSGBM sgbm;
    ImageView l{a.data(), 100, 100, 1, 0}, r{b.data(), 200, 100, 1, 0};
    EXPECT(throws<StereoException>([&] { sgbm.compute(l, r, out.data()); }));
}
static void DisparityRangeExceedsGpuLimit() {
    std::vector<uint8_t> a(64 * 64, 7);
    std::vector<int16_t> out(64 * 64);
    StereoParams p;
    p.minDisparity = 0;
    p.maxDisparity = kSgbmMaxDisparityRange + 32;
    StereoSGBM sgbm(p);
    bool ok = false;
    try {
        sgbm.compute(a.data(), a.data(), 64, 64, out.data());
    } catch (const StereoException& e) {
        ok = std::string(e.what()) ==
             "Disparity range (maxDisparity - minDisparity) exceeds GPU limit (512). Reduce the "
             "search range.";
    }
    EXPECT(ok);
}
static void InvalidRangeMessage() {
    std::vector<uint8_t> a(64 * 64, 7);
    std::vector<int16_t> out(64 * 64);
    StereoParams p;
    p.minDisparity = 10;
    p.maxDisparity = 10;
    StereoSGBM sgbm(p);
    bool ok = false;
    try {
        sgbm.compute(a.data(), a.data(), 64, 64, out.data());
    } catch (const StereoException& e) {
        ok = std::string(e.what()) ==
             "Invalid disparity range: maxDisparity must be greater than minDisparity";
    }
    EXPECT(ok);
}

// --- tests/test_sgbm.cpp:163-314 (GPU) ----------------------------------------------------
static void fillRect(std::vector<uint8_t>& img, int W, int x0, int y0, int x1, int y1, int ch,
                     const uint8_t* val) {
    for (int y = y0; y <= y1; ++y)
        for (int x = x0; x <= x1; ++x)
            for (int c = 0; c < ch; ++c) img[(size_t)(y * W + x) * ch + c] = val[c];
}
static void ComputeBasic() {
    const int W = 100, H = 100;
    std::vector<uint8_t> l(W * H, 0), r(W * H, 0);
    const uint8_t v = 128;
    fillRect(l, W, 30, 30, 70, 70, 1, &v);
    fillRect(r, W, 25, 30, 65, 70, 1, &v);
    StereoParams p;
    p.maxDisparity = 32;
    p.numDirections = 4;
    StereoSGBM sgbm(p);
    std::vector<int16_t> d(W * H, 12345);
    sgbm.compute(l.data(), r.data(), W, H, d.data());
    bool written = true;
    for (int16_t x : d) written &= x != 12345;
    EXPECT(written);
    EXPECT(sgbm.getLastComputeTimeMs() > 0.0f);
    std::printf("  SGBM compute time: %.3f ms\n", sgbm.getLastComputeTimeMs());
}
static void ColorInput() {
    const int W = 100, H = 100;
    std::vector<uint8_t> l(W * H * 3, 0), r(W * H * 3, 0);
    const uint8_t v[3] = {128, 64, 32};
    fillRect(l, W, 30, 30, 70, 70, 3, v);
    fillRect(r, W, 25, 30, 65, 70, 3, v);
    StereoParams p;
    p.maxDisparity = 32;
    StereoSGBM sgbm(p);
    std::vector<int16_t> d(W * H);
    sgbm.compute(ImageView{l.data(), W, H, 3, 0}, ImageView{r.data(), W, H, 3, 0}, d.data());
    EXPECT(true);
}
static void DifferentSizesAndDisparities() {
    StereoParams p;
    p.maxDisparity = 32;
    StereoSGBM sgbm(p);
    const int sizes[4][2] = {{64, 48}, {128, 96}, {320, 240}, {640, 480}};
    std::srand(7);
    for (auto& s : sizes) {
        std::vector<uint8_t> l((size_t)s[0] * s[1]), r((size_t)s[0] * s[1]);
        for (auto& x : l) x = std::rand() & 255;
        for (auto& x : r) x = std::rand() & 255;
        std::vector<int16_t> d((size_t)s[0] * s[1]);
        sgbm.compute(l.data(), r.data(), s[0], s[1], d.data());
        std::printf("  Size %dx%d: %.3f ms\n", s[0], s[1], sgbm.getLastComputeTimeMs());
    }
    std::vector<uint8_t> l(320 * 240), r(320 * 240);
    for (auto& x : l) x = std::rand() & 255;
    for (auto& x : r) x = std::rand() & 255;
    for (int maxDisp : {32, 64, 96, 128}) {
        StereoParams q;
        q.maxDisparity = maxDisp;
        StereoSGBM m(q);
        std::vector<int16_t> d(320 * 240);
        m.compute(l.data(), r.data(), 320, 240, d.data());
        std::printf("  maxDisparity=%d: %.3f ms (cost %.3f, aggregation %.3f, post %.3f)\n", maxDisp,
                    m.getLastComputeTimeMs(), m.getCostComputeTimeMs(), m.getAggregationTimeMs(),
                    m.getPostProcessTimeMs());
    }
    EXPECT(true);
}

// --- tests/test_sgbm.cpp:81-160 (DeviceBuffer, the reference's CudaBuffer) ----------------
static void BufferEmpty() {
    DeviceBuffer<float> buffer;
    EXPECT(buffer.empty());
    EXPECT(buffer.size() == 0);
    EXPECT(buffer.data() == nullptr);
    // Counts past the end are refused before any device call.
    float x = 0.0f;
    EXPECT(throws<StereoException>([&] { buffer.copyFrom(&x, 1); }));
    EXPECT(throws<StereoException>([&] { buffer.copyTo(&x, 1); }));
    DeviceBuffer<float> moved(std::move(buffer));
    EXPECT(moved.empty());
}
static void BufferAllocation() {
    DeviceBuffer<float> buffer(1000);
    EXPECT(buffer.size() == 1000);
    EXPECT(buffer.bytes() == 4000);
    EXPECT(!buffer.empty());
    EXPECT(buffer.data() != nullptr);
    float* before = buffer.data();
    buffer.allocate(1000);  // same size: kept
    EXPECT(buffer.data() == before);
    DeviceBuffer<float> moved(std::move(buffer));
    EXPECT(moved.data() == before && moved.size() == 1000);
    EXPECT(buffer.empty() && buffer.data() == nullptr);
    DeviceBuffer<float> assigned;
    assigned = std::move(moved);
    EXPECT(assigned.data() == before && moved.empty());
}
static void BufferResize() {
    DeviceBuffer<int> buffer(100);
    EXPECT(buffer.size() == 100);
    buffer.resize(500);
    EXPECT(buffer.size() == 500);
    buffer.resize(50);
    EXPECT(buffer.size() == 50);
    buffer.resize(0);
    EXPECT(buffer.empty());
}
static void BufferZeroMemory() {
    DeviceBuffer<int> buffer(100);
    std::vector<int> ones(100, -1);
    buffer.copyFrom(ones.data());
    buffer.zero();
    std::vector<int> host(100, 7);
    buffer.copyTo(host.data());
    bool zero = true;
    for (int v : host) zero &= v == 0;
    EXPECT(zero);
}
static void BufferCopyRoundtrip() {
    std::vector<float> original = {1.0f, 2.0f, 3.0f, 4.0f, 5.0f};
    DeviceBuffer<float> buffer(original.size());
    buffer.copyFrom(original.data());
    std::vector<float> result(original.size());
    buffer.copyTo(result.data());
    EXPECT(result == original);
    // A partial copy touches only its count.
    std::vector<float> part = {9.0f, 8.0f};
    buffer.copyFrom(part.data(), 2);
    buffer.copyTo(result.data());
    EXPECT(result[0] == 9.0f && result[1] == 8.0f && result[2] == 3.0f && result[4] == 5.0f);
    // Device to device into a larger buffer; a larger source is refused.
    DeviceBuffer<float> big = make2DBuffer<float>(4, 2);
    EXPECT(big.size() == 8);
    big.zero();
    big.copyFromDevice(buffer);
    std::vector<float> b(8);
    big.copyTo(b.data());
    EXPECT(b[0] == 9.0f && b[4] == 5.0f && b[5] == 0.0f && b[7] == 0.0f);
    EXPECT(throws<StereoException>([&] { buffer.copyFromDevice(big); }));
}
static void BufferLargeAllocation() {
    const size_t size = 100 * 1024 * 1024 / sizeof(float);
    DeviceBuffer<float> buffer(size);
    EXPECT(buffer.size() == size);
}
static void BufferTooLarge() {
    // More than any device holds: GpuException with the out-of-memory status, nothing kept.
    const size_t huge = size_t(1) << 60;
    bool ok = false;
    try {
        DeviceBuffer<uint8_t> buffer(huge);
    } catch (const GpuException& e) {
        ok = e.error_code == LUMEN_ENOMEM;
    }
    EXPECT(ok);
    DeviceBuffer<int> after(16);  // the failure leaves no error behind
    EXPECT(after.size() == 16);
}

static void DeviceImagesCompute() {
    // computeDevice on DeviceBuffers equals compute on the same host images.
    const int W = 160, H = 120;
    std::vector<uint8_t> l(W * H), r(W * H);
    std::srand(11);
    for (int i = 0; i < W * H; ++i) l[i] = static_cast<uint8_t>(std::rand() & 255);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) r[y * W + x] = l[y * W + (x + 7 < W ? x + 7 : W - 1)];
    StereoParams p;
    p.maxDisparity = 32;
    StereoSGBM sgbm(p);
    std::vector<int16_t> host(W * H), dev(W * H, 12345);
    sgbm.compute(l.data(), r.data(), W, H, host.data());
    auto dl = make2DBuffer<uint8_t>(W, H), dr = make2DBuffer<uint8_t>(W, H);
    auto dd = make2DBuffer<int16_t>(W, H);
    dl.copyFrom(l.data());
    dr.copyFrom(r.data());
    sgbm.computeDevice(dl.data(), dr.data(), W, H, dd.data());
    dd.copyTo(dev.data());
    EXPECT(dev == host);
    EXPECT(sgbm.getLastComputeTimeMs() > 0.0f);
}

// --- tests/test_rectification.cpp:21-145 (host) ------------------------------------------
static void CameraDefaultInvalid() {
    CameraParams p;
    EXPECT(!p.isValid());
}
static void CameraComputeDerivedValues() {
    CameraParams p;
    p.width = 640;
    p.height = 480;
    p.K0 = {500, 0, 320, 0, 500, 240, 0, 0, 1};
    p.K1 = p.K0;
    p.R = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    p.T = {-0.1, 0, 0};
    p.D0.assign(5, 0.0);
    p.D1.assign(5, 0.0);
    p.computeDerivedValues();
    EXPECT(std::fabs(p.baseline - 0.1) < 1e-6);
    EXPECT(std::fabs(p.focalLength - 500.0) < 1e-6);
    EXPECT(p.isValid());
}
static void CalibNonexistentFile() {
    CameraParams p;
    EXPECT(!loadMiddleburyCalib("/nonexistent/path/calib.txt", p));
}
static void CalibMiddleburyFormat() {
    const std::string path = "/tmp/lumen_test_calib.txt";
    {
        std::ofstream f(path);
        f << "cam0=[4161.221 0 1445.577; 0 4161.221 984.686; 0 0 1]\n"
             "cam1=[4161.221 0 1654.636; 0 4161.221 984.686; 0 0 1]\n"
             "doffs=209.059\nbaseline=176.252\nwidth=2880\nheight=1988\nndisp=280\n";
    }
    CameraParams p;
    EXPECT(loadMiddleburyCalib(path, p));
    EXPECT(p.width == 2880 && p.height == 1988);
    EXPECT(p.focalLength == 4161.221);
    EXPECT(std::fabs(p.baseline - 0.176252) < 1e-12);
    EXPECT(p.disparityOffset == 209 && p.numDisparities == 280);
    EXPECT(p.isValid());
    std::remove(path.c_str());
}
static void StereoParamsSaveLoad() {
    StereoParams original;
    original.minDisparity = 10;
    original.maxDisparity = 200;
    original.blockSize = 7;
    original.P1 = 15;
    original.P2 = 60;
    const std::string path = "/tmp/lumen_test_stereo_params.yaml";
    EXPECT(saveStereoParams(path, original));
    StereoParams loaded;
    EXPECT(loadStereoParams(path, loaded));
    EXPECT(loaded.minDisparity == 10 && loaded.maxDisparity == 200 && loaded.blockSize == 7);
    EXPECT(loaded.P1 == 15 && loaded.P2 == 60);
    std::remove(path.c_str());
}
static void StereoParamsMatchingCostRoundtrip() {
    StereoParams original;
    original.matchingCostMode = MatchingCostMode::Census;
    const std::string path = "/tmp/lumen_test_stereo_params_census.yaml";
    EXPECT(saveStereoParams(path, original));
    StereoParams loaded;
    EXPECT(loadStereoParams(path, loaded));
    EXPECT(loaded.matchingCostMode == MatchingCostMode::Census);
    std::remove(path.c_str());
}
// tests/test_sgbm.cpp:320-332.
static void DepthFormula() {
    const float focalLength = 1000.0f, baseline = 0.1f, disparity = 100.0f;
    EXPECT(focalLength * baseline / disparity == 1.0f);
}

// --- tests/test_rectification.cpp:160-265, tests/test_sgbm.cpp:320-349 (GPU) --------------
static void PreRectifiedPassthrough() {
    CameraParams p;
    p.width = 640;
    p.height = 480;
    StereoRectifier r(p);
    EXPECT(r.isPreRectified());
    std::vector<uint8_t> l(640 * 480 * 3), rr(640 * 480 * 3), lo(l.size()), ro(l.size());
    std::srand(3);
    for (auto& x : l) x = std::rand() & 255;
    for (auto& x : rr) x = std::rand() & 255;
    r.rectify(l.data(), rr.data(), 640, 480, 3, lo.data(), ro.data());
    EXPECT(lo == l && ro == rr);
}
static void IdentityMapExact() {
    const int W = 640, H = 480;
    std::vector<uint8_t> src(W * H), dst(W * H);
    for (auto& x : src) x = std::rand() & 255;
    std::vector<float> mx(W * H), my(W * H);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            mx[y * W + x] = static_cast<float>(x);
            my[y * W + x] = static_cast<float>(y);
        }
    rectifyImageGPU(src.data(), mx.data(), my.data(), dst.data(), W, H);
    EXPECT(dst == src);
}
static void BilinearHalfPixel() {
    const int W = 100, H = 100;
    std::vector<uint8_t> src(W * H), dst(W * H);
    std::vector<float> mx(W * H), my(W * H);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            src[y * W + x] = static_cast<uint8_t>(x * 255 / W);
            mx[y * W + x] = x + 0.5f;
            my[y * W + x] = y + 0.5f;
        }
    rectifyImageGPU(src.data(), mx.data(), my.data(), dst.data(), W, H);
    int maxDiff = 0;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            int s = 0;  // taps outside the image count as 0
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx)
                    if (x + dx < W && y + dy < H) s += src[(y + dy) * W + x + dx];
            maxDiff = std::max(maxDiff, std::abs(dst[y * W + x] - (s + 2) / 4));
        }
    EXPECT(maxDiff == 0);
}
static void DepthAndColorize() {
    const int W = 100, H = 100;
    std::vector<int16_t> d(W * H);
    for (int i = 0; i < W * H; ++i) d[i] = static_cast<int16_t>(16 * (1 + i % 200));
    std::vector<float> z(W * H);
    disparityToDepth(d.data(), z.data(), W, H, 1000.0, 0.1);
    EXPECT(z[99] == 1.0f);  // 100 px -> 1 m
    std::vector<uint8_t> rgb(W * H * 3);
    colorizeDepth(z.data(), rgb.data(), W, H, 1.0f, 10.0f);
    EXPECT(rgb.size() == static_cast<size_t>(W * H * 3));
}

int main(int argc, char** argv) {
    const bool gpu = argc > 1 && std::string(argv[1]) == "--gpu";
    std::vector<std::pair<const char*, std::function<void()>>> tests = {
        {"StereoParamsTest.DefaultValues", DefaultValues},
        {"StereoParamsTest.InvalidBlockSize", InvalidBlockSize},
        {"StereoParamsTest.InvalidDisparity", InvalidDisparity},
        {"StereoParamsTest.InvalidPenalties", InvalidPenalties},
        {"StereoParamsTest.AutoPenalties", AutoPenalties},
        {"StereoParamsTest.DefaultsMatchCAbi", DefaultsMatchCAbi},
        {"SGBMTest.EmptyInput", EmptyInput},
        {"SGBMTest.SizeMismatch", SizeMismatch},
        {"SGBMTest.DisparityRangeExceedsGpuLimit", DisparityRangeExceedsGpuLimit},
        {"SGBMTest.InvalidRangeMessage", InvalidRangeMessage},
        {"CameraParamsTest.DefaultInvalid", CameraDefaultInvalid},
        {"CameraParamsTest.ComputeDerivedValues", CameraComputeDerivedValues},
        {"CalibrationLoadTest.NonexistentFile", CalibNonexistentFile},
        {"CalibrationLoadTest.MiddleburyFormat", CalibMiddleburyFormat},
        {"ConfigTest.StereoParamsSaveLoad", StereoParamsSaveLoad},
        {"ConfigTest.StereoParamsMatchingCostRoundtrip", StereoParamsMatchingCostRoundtrip},
        {"DepthMapTest.DisparityToDepth", DepthFormula},
        {"DeviceBufferTest.EmptyBuffer", BufferEmpty},
    };
    if (gpu) {
        tests.push_back({"DeviceBufferTest.Allocation", BufferAllocation});
        tests.push_back({"DeviceBufferTest.Resize", BufferResize});
        tests.push_back({"DeviceBufferTest.ZeroMemory", BufferZeroMemory});
        tests.push_back({"DeviceBufferTest.CopyRoundtrip", BufferCopyRoundtrip});
        tests.push_back({"DeviceBufferTest.LargeAllocation", BufferLargeAllocation});
        tests.push_back({"DeviceBufferTest.TooLarge", BufferTooLarge});
        tests.push_back({"SGBMTest.DeviceImagesCompute", DeviceImagesCompute});
        tests.push_back({"SGBMTest.ComputeBasic", ComputeBasic});
        tests.push_back({"SGBMTest.ColorInput", ColorInput});
        tests.push_back({"SGBMTest.DifferentSizesAndDisparities", DifferentSizesAndDisparities});
        tests.push_back({"RectificationTest.PreRectifiedPassthrough", PreRectifiedPassthrough});
        tests.push_back({"RectificationTest.GPUvsOpenCV", IdentityMapExact});
        tests.push_back({"RectificationTest.BilinearInterpolation", BilinearHalfPixel});
        tests.push_back({"DepthMapTest.ColorizeDepth", DepthAndColorize});
    }
    for (auto& t : tests) {
        const int before = g_failed;
        std::printf("[ RUN ] %s\n", t.first);
        std::fflush(stdout);
        try {
            t.second();
        } catch (const std::exception& e) {
            std::printf("  FAILED: exception %s\n", e.what());
            ++g_failed;
        }
        std::printf("[ %s ] %s\n", g_failed == before ? " OK " : "FAIL", t.first);
    }
    std::printf("%d failure(s)\n", g_failed);
    return g_failed ? 1 : 0;
}
