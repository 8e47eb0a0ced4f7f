// evaluate.cpp - PFM ground truth and the disparity metrics of the evaluation tool. Host code.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "lumen_evaluate.h"
#include "lumen_sgbm.h"

namespace lumen {
void set_last_error(const std::string& msg);
}

namespace {

int fail(const std::string& msg) {
    lumen::set_last_error(msg);
    return LUMEN_EINVAL;
}

bool host_little_endian() {
    const uint16_t x = 1;
    uint8_t b;
    std::memcpy(&b, &x, 1);
    return b == 1;
}

void swap32(float& v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    u = (u >> 24) | ((u >> 8) & 0xFF00u) | ((u << 8) & 0xFF0000u) | (u << 24);
    std::memcpy(&v, &u, 4);
}

}  // namespace

// tools/evaluate.cpp:45-92: token "Pf", then width, height and scale as text, one whitespace
// byte, then the payload, bottom row first.
extern "C" int lumen_pfm_load(const char* path, int* width, int* height, float* data,
                              size_t capacity) {
    if (!path || !width || !height) return fail("NULL argument");
    std::ifstream file(path, std::ios::binary);
    if (!file.is_open()) return fail(std::string("Failed to open PFM: ") + path);
    std::string magic;
    file >> magic;
    if (magic != "Pf") return fail("Invalid PFM magic: " + magic + " (expected Pf)");
    int w = 0, h = 0;
    float scale = 0.0f;
    file >> w >> h >> scale;
    file.get();
    if (!file || w <= 0 || h <= 0) return fail("Invalid PFM dimensions");
    *width = w;
    *height = h;
    if (!data) return LUMEN_OK;
    const size_t n = (size_t)w * h;
    if (capacity < n) return fail("PFM buffer too small");
    std::vector<float> rows(n);
    file.read(reinterpret_cast<char*>(rows.data()), static_cast<std::streamsize>(n * sizeof(float)));
    if (!file) return fail("Failed to read PFM payload");
    const bool swap = (scale < 0.0f) != host_little_endian();
    for (int y = 0; y < h; ++y) {
        const float* src = rows.data() + (size_t)(h - 1 - y) * w;
        float* dst = data + (size_t)y * w;
        for (int x = 0; x < w; ++x) {
            float v = src[x];
            if (swap) swap32(v);
            dst[x] = v;
        }
    }
    return LUMEN_OK;
}

extern "C" int lumen_pfm_save(const char* path, const float* data, int width, int height) {
    if (!path || !data || width <= 0 || height <= 0) return fail("Invalid PFM arguments");
    std::ofstream f(path, std::ios::binary);
    if (!f.is_open()) return fail(std::string("Failed to open PFM for writing: ") + path);
    f << "Pf\n" << width << " " << height << "\n-1.0\n";
    const bool swap = !host_little_endian();
    std::vector<float> row(width);
    for (int y = height - 1; y >= 0; --y) {
        for (int x = 0; x < width; ++x) {
            row[x] = data[(size_t)y * width + x];
            if (swap) swap32(row[x]);
        }
        f.write(reinterpret_cast<const char*>(row.data()), static_cast<std::streamsize>(width * sizeof(float)));
    }
    f.close();
    return f ? LUMEN_OK : fail(std::string("Failed to write PFM: ") + path);
}

// tools/evaluate.cpp:163-221.
extern "C" int lumen_evaluate_disparity(const int16_t* disparity, const float* gt, int width,
                                        int height, lumen_eval_result* out) {
    if (!disparity || !gt || !out || width <= 0 || height <= 0) return fail("Invalid arguments");
    const size_t n = (size_t)width * height;
    std::vector<float> errors;
    int64_t bad1 = 0, bad2 = 0, bad4 = 0;
    double sumSq = 0.0, sumAbs = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const float g = gt[i];
        const float ours = static_cast<float>(disparity[i]) * (1.0f / 16.0f);
        if (g <= 0 || g > 1000 || std::isinf(g)) continue;
        if (ours <= 0) continue;
        const float e = std::fabs(ours - g);
        errors.push_back(e);
        sumAbs += e;
        sumSq += e * e;
        if (e > 1.0f) ++bad1;
        if (e > 2.0f) ++bad2;
        if (e > 4.0f) ++bad4;
    }
    const int64_t valid = static_cast<int64_t>(errors.size());
    if (valid == 0) return fail("No valid pixels to evaluate!");
    std::sort(errors.begin(), errors.end(), [](float a, float b) {
        if (std::isnan(a)) return false;
        if (std::isnan(b)) return true;
        return a < b;
    });
    lumen_eval_result r;
    r.valid = valid;
    r.valid_pct = 100.0 * valid / (double)(height * (double)width);
    r.mae = sumAbs / valid;
    r.rms = std::sqrt(sumSq / valid);
    r.bad1 = 100.0 * bad1 / valid;
    r.bad2 = 100.0 * bad2 / valid;
    r.bad4 = 100.0 * bad4 / valid;
    const size_t sz = errors.size();
    r.p50 = errors[sz * 50 / 100];
    r.p90 = errors[sz * 90 / 100];
    r.p95 = errors[sz * 95 / 100];
    r.p99 = errors[sz * 99 / 100];
    *out = r;
    return LUMEN_OK;
}
