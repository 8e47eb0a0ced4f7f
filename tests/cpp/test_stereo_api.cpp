// C++ API tests mirroring the reference's tests/test_sgbm.cpp (StereoParams, validation and
// exception cases, output shape/type on GPU). gtest is not available, so this is a minimal
// runner: `test_stereo_api` runs the host-only cases, `test_stereo_api --gpu` adds the cases
// that compute on the GPU. Exit status 0 = all passed.
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <vector>

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
    };
    if (gpu) {
        tests.push_back({"SGBMTest.ComputeBasic", ComputeBasic});
        tests.push_back({"SGBMTest.ColorInput", ColorInput});
        tests.push_back({"SGBMTest.DifferentSizesAndDisparities", DifferentSizesAndDisparities});
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
