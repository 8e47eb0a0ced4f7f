// synth_scene.cpp - deterministic synthetic stereo pairs for tests and benchmarks.
//
// Left image: N(0,1) noise blurred with a Gaussian (sigma 1.5, radius 5, mirrored borders),
// rescaled to mean 128 / std 40 (std 5 on the low-texture region), rounded and clipped.
// Ground truth: 3-6 slanted planes d = c + a(x-cx) + b(y-cy) on a Voronoi partition, plus 2-4
// fronto-parallel rectangles in front of the planes under them. Right image: the left image
// forward-warped to x - round(d) with a z-buffer (larger d wins), holes filled by a second
// noise image, then gain U[0.95,1.05], bias U[-5,5] and N(0,2) noise.
//
// Randomness comes from splitmix64 seeding a xoshiro256** generator and Box-Muller normals, so
// one seed gives the same images on every machine.
#include "lumen_sgbm.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (int i = 0; i < 4; ++i) {
            z += 0x9E3779B97F4A7C15ull;
            uint64_t t = z;
            t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ull;
            t = (t ^ (t >> 27)) * 0x94D049BB133111EBull;
            s[i] = t ^ (t >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    int range(int lo, int hi) { return lo + static_cast<int>(next() % (uint64_t)(hi - lo + 1)); }
    double gauss() {
        double u1 = 1.0 - uniform();  // (0,1]
        double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

int mirror(int i, int n) {
    if (n == 1) return 0;
    while (i < 0 || i >= n) {
        if (i < 0) i = -i;
        if (i >= n) i = 2 * (n - 1) - i;
    }
    return i;
}

// N(0,1) noise blurred by a separable Gaussian, then standardised to zero mean, unit std.
std::vector<double> band_noise(Rng& rng, int W, int H) {
    const int R = 5;
    double k[2 * R + 1], ks = 0;
    for (int i = -R; i <= R; ++i) {
        k[i + R] = std::exp(-(i * i) / (2.0 * 1.5 * 1.5));
        ks += k[i + R];
    }
    for (double& v : k) v /= ks;
    std::vector<double> a((size_t)W * H), b((size_t)W * H);
    for (double& v : a) v = rng.gauss();
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            double acc = 0;
            for (int i = -R; i <= R; ++i) acc += k[i + R] * a[(size_t)y * W + mirror(x + i, W)];
            b[(size_t)y * W + x] = acc;
        }
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            double acc = 0;
            for (int i = -R; i <= R; ++i) acc += k[i + R] * b[(size_t)mirror(y + i, H) * W + x];
            a[(size_t)y * W + x] = acc;
        }
    double mean = 0, sq = 0;
    for (double v : a) mean += v;
    mean /= a.size();
    for (double v : a) sq += (v - mean) * (v - mean);
    const double sd = std::sqrt(sq / a.size());
    for (double& v : a) v = (v - mean) / (sd > 0 ? sd : 1.0);
    return a;
}

uint8_t to_u8(double v) {
    const double r = std::floor(v + 0.5);
    return static_cast<uint8_t>(std::min(255.0, std::max(0.0, r)));
}

}  // namespace

extern "C" int lumen_synth_pair(int W, int H, float dmin, float dmax, uint64_t seed,
                                float lowtex_frac, uint8_t* left, uint8_t* right, float* gt) {
    if (W <= 0 || H <= 0 || !left || !right || !(dmax >= dmin)) return LUMEN_EINVAL;
    Rng rng(seed);
    const size_t n = (size_t)W * H;
    std::vector<double> tex = band_noise(rng, W, H);

    // Low-texture region: one rectangle covering lowtex_frac of the area.
    int lx0 = 0, ly0 = 0, lx1 = 0, ly1 = 0;
    if (lowtex_frac > 0) {
        const double f = std::min(1.0, (double)lowtex_frac);
        const double aspect = rng.uniform(0.6, 1.6);
        int rw = std::min(W, std::max(1, (int)std::lround(std::sqrt(f * W * H * aspect))));
        int rh = std::min(H, std::max(1, (int)std::lround(f * W * H / rw)));
        lx0 = rng.range(0, W - rw);
        ly0 = rng.range(0, H - rh);
        lx1 = lx0 + rw;
        ly1 = ly0 + rh;
    }
    std::vector<uint8_t> L(n);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const bool low = x >= lx0 && x < lx1 && y >= ly0 && y < ly1;
            L[(size_t)y * W + x] = to_u8(128.0 + (low ? 5.0 : 40.0) * tex[(size_t)y * W + x]);
        }

    // Ground truth: Voronoi planes.
    const double span = std::max(0.0, (double)dmax - dmin);
    const double planeMax = dmin + 0.8 * span;
    const int nplanes = rng.range(3, 6);
    struct Plane {
        double cx, cy, c, a, b;
    };
    std::vector<Plane> planes(nplanes);
    for (Plane& p : planes) {
        p.cx = rng.uniform(0, W);
        p.cy = rng.uniform(0, H);
        p.c = rng.uniform(dmin + 0.1 * span, dmin + 0.6 * span);
        p.a = rng.uniform(-0.3, 0.3) * span / std::max(1, W);
        p.b = rng.uniform(-0.3, 0.3) * span / std::max(1, H);
    }
    std::vector<float> G(n);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            int best = 0;
            double bd = 1e300;
            for (int i = 0; i < nplanes; ++i) {
                const double dx = x - planes[i].cx, dy = y - planes[i].cy;
                const double dd = dx * dx + dy * dy;
                if (dd < bd) {
                    bd = dd;
                    best = i;
                }
            }
            const Plane& p = planes[best];
            double d = p.c + p.a * (x - p.cx) + p.b * (y - p.cy);
            d = std::min(planeMax, std::max((double)dmin, d));
            G[(size_t)y * W + x] = static_cast<float>(d);
        }
    // Fronto-parallel occluders above the planes they cover.
    const int nocc = rng.range(2, 4);
    for (int i = 0; i < nocc; ++i) {
        const int rw = std::max(1, (int)std::lround(rng.uniform(0.08, 0.25) * W));
        const int rh = std::max(1, (int)std::lround(rng.uniform(0.08, 0.25) * H));
        const int x0 = rng.range(0, W - rw), y0 = rng.range(0, H - rh);
        float under = (float)dmin;
        for (int y = y0; y < y0 + rh; ++y)
            for (int x = x0; x < x0 + rw; ++x) under = std::max(under, G[(size_t)y * W + x]);
        const double lift = rng.uniform(2.0, 2.0 + 0.1 * span);
        const float d = static_cast<float>(std::min((double)dmax, under + lift));
        for (int y = y0; y < y0 + rh; ++y)
            for (int x = x0; x < x0 + rw; ++x) G[(size_t)y * W + x] = d;
    }

    // Right image: forward warp with z-buffer, holes from fresh noise, photometric change.
    std::vector<double> fill = band_noise(rng, W, H);
    std::vector<float> zb(n, -1e30f);
    std::vector<int> src(n, -1);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const float d = G[(size_t)y * W + x];
            const int xr = x - (int)std::lround(d);
            if (xr < 0 || xr >= W) continue;
            const size_t q = (size_t)y * W + xr;
            if (d > zb[q]) {
                zb[q] = d;
                src[q] = x;
            }
        }
    const double gain = rng.uniform(0.95, 1.05), bias = rng.uniform(-5.0, 5.0);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t q = (size_t)y * W + x;
            const double v = src[q] >= 0 ? (double)L[(size_t)y * W + src[q]]
                                         : 128.0 + 40.0 * fill[q];
            right[q] = to_u8(gain * v + bias + 2.0 * rng.gauss());
        }
    std::copy(L.begin(), L.end(), left);
    if (gt) std::copy(G.begin(), G.end(), gt);
    return LUMEN_OK;
}
