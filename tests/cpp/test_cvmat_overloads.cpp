// The library's optional cv::Mat overloads (include/stereo/*.h under LUMEN_WITH_OPENCV),
// compiled and run against the cv::Mat test double in tests/cpp/mock_opencv (OpenCV is not
// installed here). Each overload must give what its raw-pointer form gives on the same pixels,
// including for matrices that are not continuous (ROIs, whose rows are further apart than their
// width), and must raise the reference's messages for empty and mismatched inputs
// (src/sgbm_gpu.cu:148-166). `test_cvmat_overloads` runs the host-only cases,
// `test_cvmat_overloads --gpu` adds the cases that compute. Exit status 0 = all passed.
#define LUMEN_WITH_OPENCV 1
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <utility>
#include <vector>

#include "stereo/depth_map.h"
#include "stereo/rectification.h"
#include "stereo/sgbm_gpu.h"

using namespace stereo;

static int g_failed = 0;
#define EXPECT(cond)                                                        \
    do {                                                                    \
        if (!(cond)) {                                                      \
            std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            ++g_failed;                                                     \
        }                                                                   \
    } while (0)

template <typename F>
static std::string message_of(F&& f) {
    try {
        f();
    } catch (const StereoException& e) {
        return e.what();
    } catch (...) {
        return "other exception";
    }
    return "";
}

// A textured pair: left random, right = left shifted by `shift` pixels (edge repeated).
static void make_pair(cv::Mat& l, cv::Mat& r, int W, int H, int C, int shift, unsigned seed) {
    l.create(H, W, C == 3 ? CV_8UC3 : CV_8UC1);
    r.create(H, W, C == 3 ? CV_8UC3 : CV_8UC1);
    std::srand(seed);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W * C; ++x) l.data[y * l.step + x] = static_cast<uint8_t>(std::rand());
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const int sx = x + shift < W ? x + shift : W - 1;
            for (int c = 0; c < C; ++c) r.data[y * r.step + x * C + c] = l.data[y * l.step + sx * C + c];
        }
}

static bool same_disparity(const cv::Mat& m, const std::vector<int16_t>& v) {
    if (m.type() != CV_16SC1 || m.rows * m.cols != static_cast<int>(v.size())) return false;
    for (int y = 0; y < m.rows; ++y)
        for (int x = 0; x < m.cols; ++x)
            if (m.at<int16_t>(y, x) != v[static_cast<size_t>(y) * m.cols + x]) return false;
    return true;
}

// --- host cases: the reference's validation messages (src/sgbm_gpu.cu:148-166) -------------
static void EmptyMat() {
    StereoSGBM sgbm;
    cv::Mat a, b, d;
    EXPECT(message_of([&] { sgbm.compute(a, b, d); }) == "Input images are empty");
}
static void SizeMismatchMat() {
    StereoSGBM sgbm;
    cv::Mat a(100, 100, CV_8UC1), b(100, 200, CV_8UC1), d;
    EXPECT(message_of([&] { sgbm.compute(a, b, d); }) == "Left and right images must have same size");
}
static void MockSemantics() {
    // The double behaves as the overloads expect: create keeps a matching allocation, an ROI
    // shares data with its parent's step, clone is continuous.
    cv::Mat m(10, 20, CV_16SC1);
    uint8_t* p = m.data;
    m.create(10, 20, CV_16SC1);
    EXPECT(m.data == p && m.step == 40 && m.isContinuous());
    cv::Mat roi(m, cv::Rect(2, 3, 5, 4));
    EXPECT(!roi.isContinuous() && roi.step == 40 && roi.data == m.data + 3 * 40 + 2 * 2);
    EXPECT(roi.clone().isContinuous() && roi.clone().step == 10);
}

// --- GPU cases ------------------------------------------------------------------------------
static void ComputeMatEqualsPointerForm() {
    const int W = 192, H = 96;
    cv::Mat l, r;
    make_pair(l, r, W, H, 1, 9, 3);
    StereoParams p;
    p.maxDisparity = 48;
    StereoSGBM sgbm(p);
    std::vector<int16_t> want(static_cast<size_t>(W) * H);
    sgbm.compute(l.data, r.data, W, H, want.data());
    cv::Mat d;
    sgbm.compute(l, r, d);
    EXPECT(same_disparity(d, want));
    EXPECT(sgbm.getLastComputeTimeMs() > 0.0f);
}
static void ComputeMatRoi() {
    // Inputs that are windows of larger images (rows further apart than their width), and an
    // output ROI: the overload passes each step through, so the result equals the packed form.
    const int W = 160, H = 80;
    cv::Mat bigL, bigR;
    make_pair(bigL, bigR, W + 37, H + 11, 1, 6, 5);
    cv::Mat l(bigL, cv::Rect(13, 5, W, H)), r(bigR, cv::Rect(13, 5, W, H));
    StereoParams p;
    p.maxDisparity = 32;
    p.numDirections = 8;
    p.matchingCostMode = MatchingCostMode::Census;
    StereoSGBM sgbm(p);
    const cv::Mat lp = l.clone(), rp = r.clone();
    std::vector<int16_t> want(static_cast<size_t>(W) * H);
    sgbm.compute(lp.data, rp.data, W, H, want.data());
    cv::Mat d;
    sgbm.compute(l, r, d);
    EXPECT(same_disparity(d, want));
    cv::Mat bigOut(H + 4, W + 9, CV_16SC1);
    cv::Mat dr(bigOut, cv::Rect(4, 2, W, H));
    sgbm.compute(l, r, dr);  // already the right size and type: written in place, with its step
    EXPECT(dr.data == bigOut.data + 2 * bigOut.step + 4 * 2);
    EXPECT(same_disparity(dr.clone(), want));
}
static void ComputeMatBgr() {
    const int W = 128, H = 64;
    cv::Mat l, r;
    make_pair(l, r, W, H, 3, 5, 7);
    StereoParams p;
    p.maxDisparity = 32;
    StereoSGBM sgbm(p);
    std::vector<int16_t> want(static_cast<size_t>(W) * H);
    sgbm.compute(ImageView{l.data, W, H, 3, l.step}, ImageView{r.data, W, H, 3, r.step}, want.data());
    cv::Mat d;
    sgbm.compute(l, r, d);
    EXPECT(same_disparity(d, want));
}
static void DepthAndColourMat() {
    const int W = 64, H = 40;
    cv::Mat big(H + 2, W + 6, CV_16SC1);
    for (int y = 0; y < big.rows; ++y)
        for (int x = 0; x < big.cols; ++x) big.at<int16_t>(y, x) = static_cast<int16_t>(16 * (1 + (x * 7 + y) % 90));
    cv::Mat disp(big, cv::Rect(3, 1, W, H));  // not continuous
    const cv::Mat packed = disp.clone();
    std::vector<float> want(static_cast<size_t>(W) * H);
    disparityToDepth(reinterpret_cast<const int16_t*>(packed.data), want.data(), W, H, 1000.0, 0.1);
    cv::Mat depth;
    disparityToDepth(disp, depth, 1000.0, 0.1);
    bool ok = depth.type() == CV_32FC1 && depth.rows == H && depth.cols == W;
    for (int i = 0; ok && i < W * H; ++i) ok = reinterpret_cast<const float*>(depth.data)[i] == want[i];
    EXPECT(ok);
    std::vector<uint8_t> rgb(static_cast<size_t>(W) * H * 3);
    colorizeDepth(want.data(), rgb.data(), W, H, 1.0f, 10.0f);
    cv::Mat col;
    colorizeDepth(depth, col, 1.0f, 10.0f);
    ok = col.type() == CV_8UC3 && col.rows == H && col.cols == W;
    for (size_t i = 0; ok && i < rgb.size(); ++i) ok = col.data[i] == rgb[i];
    EXPECT(ok);
}
static void RemapMat() {
    // Identity maps reproduce the source exactly (the reference's GPUvsOpenCV case), also for a
    // source that is an ROI.
    const int W = 90, H = 50;
    cv::Mat big(H + 3, W + 5, CV_8UC1);
    for (int y = 0; y < big.rows; ++y)
        for (int x = 0; x < big.cols; ++x) big.data[y * big.step + x] = static_cast<uint8_t>(x * 3 + y * 11);
    cv::Mat src(big, cv::Rect(2, 1, W, H));
    cv::Mat mx(H, W, CV_32FC1), my(H, W, CV_32FC1);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            mx.at<float>(y, x) = static_cast<float>(x);
            my.at<float>(y, x) = static_cast<float>(y);
        }
    cv::Mat dst;
    rectifyImageGPU(src, mx, my, dst);
    bool ok = dst.rows == H && dst.cols == W && dst.type() == CV_8UC1;
    for (int y = 0; ok && y < H; ++y)
        for (int x = 0; ok && x < W; ++x) ok = dst.data[y * dst.step + x] == src.data[y * src.step + x];
    EXPECT(ok);
}

int main(int argc, char** argv) {
    const bool gpu = argc > 1 && std::string(argv[1]) == "--gpu";
    std::vector<std::pair<const char*, std::function<void()>>> tests = {
        {"CvMat.EmptyInput", EmptyMat},
        {"CvMat.SizeMismatch", SizeMismatchMat},
        {"CvMat.TestDoubleSemantics", MockSemantics},
    };
    if (gpu) {
        tests.push_back({"CvMat.ComputeEqualsPointerForm", ComputeMatEqualsPointerForm});
        tests.push_back({"CvMat.ComputeRoiInputsAndOutput", ComputeMatRoi});
        tests.push_back({"CvMat.ComputeBgr", ComputeMatBgr});
        tests.push_back({"CvMat.DepthAndColour", DepthAndColourMat});
        tests.push_back({"CvMat.RemapIdentity", RemapMat});
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
