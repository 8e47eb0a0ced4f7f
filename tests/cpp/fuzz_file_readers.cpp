// Robustness of the host-side file readers (parameter YAML, Middlebury calib.txt, PFM): 20000
// random files built from the formats' own tokens plus random bytes must be handled without a
// memory error. Built and run under AddressSanitizer/UBSan by tests/test_host_sanitizers.py.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <string>
#include "lumen_config.h"
#include "lumen_evaluate.h"
namespace lumen { void set_last_error(const std::string&) {} }
int main() {
    std::mt19937 rng(5);
    const char* path = "/tmp/lumen_asan_cfg.txt";
    const std::string pieces[] = {"minDisparity", "maxDisparity", ":", " ", "\n", "#", "'", "\"", "0x", "0o", "12", "-", "+",
                                  "useGPU", "true", "matchingCostMode", "Census", "[", "{", "\t", "---", "cam0=[", "]", ";",
                                  "width=", "height=", "ndisp=", "doffs=", "baseline=", "1e309", "abc", "Pf", "-1.0", "3 2"};
    for (int it = 0; it < 20000; ++it) {
        std::string s;
        const int n = rng() % 40;
        for (int k = 0; k < n; ++k) s += pieces[rng() % (sizeof(pieces) / sizeof(pieces[0]))];
        if (rng() % 4 == 0) for (int k = 0; k < 20; ++k) s += char(rng() % 256);
        { std::ofstream f(path, std::ios::binary); f << s; }
        lumen_stereo_params p{}; p.blockSize = 5; p.maxDisparity = 64; p.P1 = 1; p.P2 = 2;
        lumen_params_load_yaml(path, &p);
        lumen_camera_params c{};
        lumen_middlebury_calib_load(path, &c);
        int w = 0, h = 0;
        float buf[64];
        lumen_pfm_load(path, &w, &h, buf, 64);
    }
    lumen_stereo_params p{};
    lumen_params_save_yaml(path, &p);
    int16_t d[6] = {16, 0, -1, 32767, 5, 100};
    float g[6] = {1, 2, 1e30f, 0, -1, 6};
    lumen_eval_result r;
    lumen_evaluate_disparity(d, g, 3, 2, &r);
    std::printf("ok\n");
    return 0;
}
