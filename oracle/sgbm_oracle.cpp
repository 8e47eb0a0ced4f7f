/*
 * sgbm_oracle.cpp - CPU restatement of the reference SGBM matcher (test infrastructure only).
 *
 * The oracle is the checker: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * are its only users. Every function cites the reference lines it restates. Integer stores to
 * uint16_t are truncations mod 2^16 exactly where the reference stores into uint16_t.
 */
#include "sgbm_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace {

inline uint16_t t16(int v) { return static_cast<uint16_t>(static_cast<unsigned>(v) & 0xFFFFu); }

/* float -> int16 as the reference's device conversion does it: truncate toward zero and clamp
 * to the int16 range (src/sgbm_kernels.cu:804,855). */
inline int16_t f2i16(float v) {
    float t = std::trunc(v);
    if (!(t == t)) return 0;
    if (t > 32767.0f) return 32767;
    if (t < -32768.0f) return -32768;
    return static_cast<int16_t>(static_cast<int>(t));
}

/* Parabola refinement (src/sgbm_kernels.cu:789-804 and :836-855): all float operations in
 * IEEE single precision, round to nearest. */
inline float parabola(float bestD, uint16_t a0, uint16_t a1, uint16_t a2) {
    float c0 = static_cast<float>(a0);
    float c1 = static_cast<float>(a1);
    float c2 = static_cast<float>(a2);
    volatile float denom = 2.0f * (c0 + c2 - 2.0f * c1);
    float s = bestD;
    if (std::fabs(denom) > 1e-6f) {
        volatile float delta = (c0 - c2) / denom;
        float dl = std::fmax(-1.0f, std::fmin(1.0f, static_cast<float>(delta)));
        s += dl;
    }
    return s;
}

struct Dir {
    int dx, dy;
};
/* Path directions in the reference's launch order (src/sgbm_gpu.cu:321-421):
 * LR, RL, TB, BT, TLBR, BRTL, TRBL, BLTR; (dx,dy) is the step from predecessor to pixel. */
const Dir kDirs[8] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {-1, -1}, {-1, 1}, {1, -1}};

/* P2 reduced at intensity edges (src/sgbm_kernels.cu:304-310). */
inline int adaptive_p2(int P1, int P2, int cur, int prev) {
    int g = std::abs(cur - prev);
    return std::max(P1 + 1, P2 * 8 / std::max(8, g));
}

}  // namespace

extern "C" {

/* src/sgbm_kernels.cu:38-59; cap = max(1, preFilterCap) at src/sgbm_gpu.cu:227. */
void oracle_prefilter(const uint8_t* img, int W, int H, int preFilterCap, uint8_t* out) {
    const int cap = std::max(1, preFilterCap);
    for (int y = 0; y < H; ++y) {
        const uint8_t* row = img + static_cast<size_t>(y) * W;
        for (int x = 0; x < W; ++x) {
            int g;
            if (x > 0 && x < W - 1)
                g = int(row[x + 1]) - int(row[x - 1]);
            else if (x == 0)
                g = int(row[1]) - int(row[0]);
            else
                g = int(row[x]) - int(row[x - 1]);
            g = std::min(cap, std::max(-cap, g));
            out[static_cast<size_t>(y) * W + x] = static_cast<uint8_t>(g + cap);
        }
    }
}

/* src/sgbm_kernels.cu:217-245: bit k over the 5x5 neighbours in row-major order (centre
 * skipped) is set when neighbour >= centre; 0 within 2 px of the border. */
void oracle_census(const uint8_t* img, int W, int H, uint32_t* out) {
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            uint32_t code = 0;
            if (x >= 2 && x < W - 2 && y >= 2 && y < H - 2) {
                const uint8_t c = img[static_cast<size_t>(y) * W + x];
                int k = 0;
                for (int dy = -2; dy <= 2; ++dy)
                    for (int dx = -2; dx <= 2; ++dx) {
                        if (dx == 0 && dy == 0) continue;
                        if (img[static_cast<size_t>(y + dy) * W + (x + dx)] >= c) code |= 1u << k;
                        ++k;
                    }
            }
            out[static_cast<size_t>(y) * W + x] = code;
        }
    }
}

/* src/sgbm_kernels.cu:172-212 (the naive kernel; the shared-memory kernel :70-167 gives the
 * same values because its right tile always holds the addressed columns). Rows y0..y1-1. */
static void cost_sad_rows(const uint8_t* fl, const uint8_t* fr, int W, int H, int D, int m,
                          int blockSize, uint8_t* cost, int y0, int y1) {
    const int r = blockSize / 2;
    for (int y = y0; y < y1; ++y)
        for (int x = 0; x < W; ++x)
            for (int d = 0; d < D; ++d) {
                const int a = m + d;
                int sum = 0, count = 0;
                for (int dy = -r; dy <= r; ++dy)
                    for (int dx = -r; dx <= r; ++dx) {
                        const int lx = x + dx, rx = x + dx - a, ly = y + dy;
                        if (lx >= 0 && lx < W && ly >= 0 && ly < H && rx >= 0 && rx < W) {
                            sum += std::abs(int(fl[static_cast<size_t>(ly) * W + lx]) -
                                            int(fr[static_cast<size_t>(ly) * W + rx]));
                            ++count;
                        }
                    }
                cost[(static_cast<size_t>(y) * W + x) * D + d] =
                    static_cast<uint8_t>(count > 0 ? sum / count : 255);
            }
}

void oracle_cost_sad(const uint8_t* fl, const uint8_t* fr, int W, int H, int D, int m,
                     int blockSize, uint8_t* cost) {
    cost_sad_rows(fl, fr, W, H, D, m, blockSize, cost, 0, H);
}

/* src/sgbm_kernels.cu:250-286. Rows y0..y1-1. */
static void cost_census_rows(const uint32_t* cl, const uint32_t* cr, int W, int H, int D, int m,
                             uint8_t* cost, int y0, int y1) {
    for (int y = y0; y < y1; ++y)
        for (int x = 0; x < W; ++x) {
            uint8_t* out = cost + (static_cast<size_t>(y) * W + x) * D;
            const bool leftOk = x >= 2 && x < W - 2 && y >= 2 && y < H - 2;
            const uint32_t a = cl[static_cast<size_t>(y) * W + x];
            for (int d = 0; d < D; ++d) {
                const int rx = x - (m + d);
                if (!leftOk || !(rx >= 2 && rx < W - 2)) {
                    out[d] = 255;
                    continue;
                }
                const int h = __builtin_popcount(a ^ cr[static_cast<size_t>(y) * W + rx]);
                out[d] = static_cast<uint8_t>((h * 255 + 12) / 24);
            }
        }
}

void oracle_cost_census(const uint32_t* cl, const uint32_t* cr, int W, int H, int D, int m,
                        uint8_t* cost) {
    cost_census_rows(cl, cr, W, H, D, m, cost, 0, H);
}

/* src/sgbm_gpu.cu:312-317: SAD uses blockSize^2, every other mode the census formula. */
int oracle_cost_scale(int matchingCostMode, int blockSize, int preFilterCap) {
    if (matchingCostMode == 0) return blockSize * blockSize;
    return std::max(1, (blockSize * blockSize * 2 * preFilterCap) / 255);
}

/* src/sgbm_gpu.cu:321-421: four paths always, four more when numDirections >= 8. */
int oracle_num_paths(int numDirections) { return numDirections >= 8 ? 8 : 4; }

/* One SGM path (the recurrence of src/sgbm_kernels.cu:338-360, identical in the horizontal
 * :315-412, vertical :417-514 and diagonal :519-742 kernels). A pixel whose predecessor lies
 * outside the image starts the path with L = t16(C*costScale); otherwise
 *   L(p,d) = t16( t16(C*cs) + min(L(q,d), t16(L(q,d-1)+P1), t16(L(q,d+1)+P1),
 *                                 t16(M + P2adj)) - M ),  M = min_d L(q,d),
 * with out-of-range neighbours replaced by 65535. Pixels are visited so that every
 * predecessor is finished first, which is what the reference's launch order guarantees. */
void oracle_aggregate_path(const uint8_t* cost, const uint8_t* img, int W, int H, int D,
                           int P1, int P2, int costScale, int path, uint16_t* L) {
    const Dir dir = kDirs[path];
    const int y0 = dir.dy >= 0 ? 0 : H - 1, ystep = dir.dy >= 0 ? 1 : -1;
    const int x0 = dir.dx >= 0 ? 0 : W - 1, xstep = dir.dx >= 0 ? 1 : -1;
    for (int yi = 0; yi < H; ++yi) {
        const int y = y0 + yi * ystep;
        for (int xi = 0; xi < W; ++xi) {
            const int x = x0 + xi * xstep;
            const size_t p = static_cast<size_t>(y) * W + x;
            const uint8_t* c = cost + p * D;
            uint16_t* out = L + p * D;
            const int qx = x - dir.dx, qy = y - dir.dy;
            if (qx < 0 || qx >= W || qy < 0 || qy >= H) {
                for (int d = 0; d < D; ++d) out[d] = t16(int(c[d]) * costScale);
                continue;
            }
            const size_t q = static_cast<size_t>(qy) * W + qx;
            const uint16_t* prev = L + q * D;
            uint16_t M = 65535;
            for (int d = 0; d < D; ++d) M = std::min(M, prev[d]);
            const int P2adj = adaptive_p2(P1, P2, img[p], img[q]);
            const uint16_t c3 = t16(int(M) + P2adj);
            for (int d = 0; d < D; ++d) {
                const uint16_t Chat = t16(int(c[d]) * costScale);
                const uint16_t c0 = prev[d];
                const uint16_t c1 = d > 0 ? t16(int(prev[d - 1]) + P1) : uint16_t(65535);
                const uint16_t c2 = d < D - 1 ? t16(int(prev[d + 1]) + P1) : uint16_t(65535);
                const uint16_t mn = std::min(std::min(c0, c1), std::min(c2, c3));
                out[d] = t16(int(Chat) + int(mn) - int(M));
            }
        }
    }
}

/* src/sgbm_kernels.cu:748-805: first minimum with strict '<' from 65535, second minimum of
 * the multiset, uniqueness in uint32 arithmetic, parabola for 0 < bestD < D-1. Pixels
 * p0..p1-1. */
static void wta_left_pixels(const uint16_t* S, int D, int m, int uniquenessRatio, int16_t* disp,
                            int16_t* wta_int, size_t p0, size_t p1) {
    for (size_t p = p0; p < p1; ++p) {
        const uint16_t* s = S + p * D;
        uint16_t best = 65535, second = 65535;
        int bestD = 0;
        for (int d = 0; d < D; ++d) {
            if (s[d] < best) {
                second = best;
                best = s[d];
                bestD = d;
            } else if (s[d] < second) {
                second = s[d];
            }
        }
        if (wta_int) wta_int[p] = static_cast<int16_t>(bestD);
        if (uniquenessRatio > 0 && second != 65535) {
            const uint32_t b = best, sc = second;
            const uint32_t thresh = (b * static_cast<uint32_t>(uniquenessRatio)) / 100u;
            if (sc <= b + thresh) {
                disp[p] = -1;
                continue;
            }
        }
        float sub = static_cast<float>(bestD);
        if (bestD > 0 && bestD < D - 1) sub = parabola(sub, s[bestD - 1], s[bestD], s[bestD + 1]);
        disp[p] = f2i16((sub + static_cast<float>(m)) * 16.0f);
    }
}

void oracle_wta_left(const uint16_t* S, int W, int H, int D, int m, int uniquenessRatio,
                     int16_t* disp, int16_t* wta_int) {
    wta_left_pixels(S, D, m, uniquenessRatio, disp, wta_int, 0, static_cast<size_t>(W) * H);
}

/* src/sgbm_kernels.cu:807-856: right pixel x scans S(y, x+m+d, d) while x+m+d < W; the
 * parabola uses the three costs of that same left pixel. Requires m >= 0. Rows y0..y1-1. */
static void wta_right_rows(const uint16_t* S, int W, int D, int m, int16_t* dispR, int y0,
                           int y1) {
    for (int y = y0; y < y1; ++y)
        for (int x = 0; x < W; ++x) {
            uint16_t best = 65535;
            int bestD = 0;
            for (int d = 0; d < D; ++d) {
                const int lx = x + m + d;
                if (lx >= W) break;
                const uint16_t c = S[(static_cast<size_t>(y) * W + lx) * D + d];
                if (c < best) {
                    best = c;
                    bestD = d;
                }
            }
            float sub = static_cast<float>(bestD);
            if (bestD > 0 && bestD < D - 1) {
                const int lx = x + m + bestD;
                if (lx < W) {
                    const uint16_t* s = S + (static_cast<size_t>(y) * W + lx) * D;
                    sub = parabola(sub, s[bestD - 1], s[bestD], s[bestD + 1]);
                }
            }
            dispR[static_cast<size_t>(y) * W + x] = f2i16((sub + static_cast<float>(m)) * 16.0f);
        }
}

void oracle_wta_right(const uint16_t* S, int W, int H, int D, int m, int16_t* dispR) {
    wta_right_rows(S, W, D, m, dispR, 0, H);
}

/* src/sgbm_kernels.cu:862-890 (in place; each pixel reads only itself and dispR). */
void oracle_lr_check(int16_t* dispL, const int16_t* dispR, int W, int H, int maxDiff) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            int16_t& dl = dispL[static_cast<size_t>(y) * W + x];
            if (dl <= 0) continue;
            const int xr = x - dl / 16;
            if (xr < 0 || xr >= W) {
                dl = -1;
                continue;
            }
            const int16_t dr = dispR[static_cast<size_t>(y) * W + xr];
            if (dr <= 0 || std::abs(int(dl) - int(dr)) > maxDiff) dl = -1;
        }
}

/* src/sgbm_kernels.cu:979-1013 restated with snapshot semantics: every pixel reads the input
 * as it was before the pass. (The reference updates in place while neighbours are read, so its
 * result is not deterministic; parity is unpinned for this stage.) */
void oracle_speckle(int16_t* disp, int W, int H, int minSimilar, int maxDiff) {
    std::vector<int16_t> snap(disp, disp + static_cast<size_t>(W) * H);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const int16_t c = snap[static_cast<size_t>(y) * W + x];
            if (c <= 0) continue;
            int n = 0;
            for (int dy = -2; dy <= 2; ++dy)
                for (int dx = -2; dx <= 2; ++dx) {
                    const int nx = x + dx, ny = y + dy;
                    if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
                    const int16_t v = snap[static_cast<size_t>(ny) * W + nx];
                    if (v > 0 && std::abs(int(c) - int(v)) <= maxDiff) ++n;
                }
            if (n < minSimilar) disp[static_cast<size_t>(y) * W + x] = -1;
        }
}

/* src/sgbm_kernels.cu:892-935: upper median of the positive values in the in-bounds 3x3
 * window; the centre is kept when none is positive. */
void oracle_median(const int16_t* in, int16_t* out, int W, int H) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            int16_t v[9];
            int n = 0;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int nx = x + dx, ny = y + dy;
                    if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
                    const int16_t t = in[static_cast<size_t>(ny) * W + nx];
                    if (t > 0) v[n++] = t;
                }
            const size_t p = static_cast<size_t>(y) * W + x;
            if (n == 0) {
                out[p] = in[p];
                continue;
            }
            std::sort(v, v + n);
            out[p] = v[n / 2];
        }
}

}  // extern "C"

namespace {

/* Runs fn(lo, hi) over [0, n) split into contiguous bands on up to nthreads threads. */
template <typename F>
void parallel_bands(size_t n, int nthreads, F fn) {
    if (nthreads <= 1 || n < 2) {
        fn(size_t(0), n);
        return;
    }
    const size_t parts = std::min<size_t>(static_cast<size_t>(nthreads) * 4, n);
    std::vector<std::thread> pool;
    std::atomic<size_t> next{0};
    for (int t = 0; t < nthreads; ++t)
        pool.emplace_back([&] {
            for (size_t k; (k = next.fetch_add(1)) < parts;)
                fn(n * k / parts, n * (k + 1) / parts);
        });
    for (auto& th : pool) th.join();
}

/* Stage storage: the caller's dump buffers where given, otherwise owned vectors. */
struct Pipeline {
    const oracle_params* p;
    int W, H, D, m, npaths, cs, threads;
    std::vector<uint8_t> ownCost;
    std::vector<uint16_t> ownPaths, ownSum;
    uint8_t* cost = nullptr;
    uint16_t* paths = nullptr;
    uint16_t* sum = nullptr;
};

/* Validation of src/sgbm_gpu.cu:148-166 (messages are the product's business). Negative
 * minDisparity is outside the contract: the reference's right WTA then reads outside the
 * cost volume at y = 0. */
bool valid_args(const oracle_params* p, const uint8_t* l, const uint8_t* r, int W, int H) {
    if (!p || !l || !r || W <= 0 || H <= 0) return false;
    const int D = p->maxDisparity - p->minDisparity;
    if (D < 1 || D > 512) return false;
    if (p->minDisparity < 0) return false;
    if (p->matchingCostMode != 1 && W < 2) return false;
    return true;
}

/* Cost stage of src/sgbm_gpu.cu:197-295: census on the raw images when the mode is Census,
 * otherwise Sobel-x prefilter then SAD. */
void cost_stage(Pipeline& st, const uint8_t* left, const uint8_t* right) {
    const size_t n = static_cast<size_t>(st.W) * st.H;
    if (st.p->matchingCostMode == 1) {
        std::vector<uint32_t> cl(n), cr(n);
        oracle_census(left, st.W, st.H, cl.data());
        oracle_census(right, st.W, st.H, cr.data());
        parallel_bands(st.H, st.threads, [&](size_t y0, size_t y1) {
            cost_census_rows(cl.data(), cr.data(), st.W, st.H, st.D, st.m, st.cost,
                             static_cast<int>(y0), static_cast<int>(y1));
        });
    } else {
        std::vector<uint8_t> fl(n), fr(n);
        oracle_prefilter(left, st.W, st.H, st.p->preFilterCap, fl.data());
        oracle_prefilter(right, st.W, st.H, st.p->preFilterCap, fr.data());
        parallel_bands(st.H, st.threads, [&](size_t y0, size_t y1) {
            cost_sad_rows(fl.data(), fr.data(), st.W, st.H, st.D, st.m, st.p->blockSize,
                          st.cost, static_cast<int>(y0), static_cast<int>(y1));
        });
    }
}

/* The paths of src/sgbm_gpu.cu:321-421, each independent of the others (one per thread). */
void path_stage(Pipeline& st, const uint8_t* left) {
    const size_t vol = static_cast<size_t>(st.W) * st.H * st.D;
    parallel_bands(st.npaths, std::min(st.threads, st.npaths), [&](size_t r0, size_t r1) {
        for (size_t r = r0; r < r1; ++r)
            oracle_aggregate_path(st.cost, left, st.W, st.H, st.D, st.p->P1, st.p->P2, st.cs,
                                  static_cast<int>(r), st.paths + r * vol);
    });
}

/* Path sum (the uint16 accumulation of aggCost, src/sgbm_gpu.cu:321-421), WTA and
 * post-processing of src/sgbm_gpu.cu:432-515. */
void finish(Pipeline& st, int16_t* out, const oracle_dumps* dumps) {
    const size_t n = static_cast<size_t>(st.W) * st.H;
    const size_t vol = n * st.D;
    uint16_t* S = st.sum;
    parallel_bands(vol, st.threads, [&](size_t i0, size_t i1) {
        for (size_t i = i0; i < i1; ++i) {
            int acc = 0;
            for (int r = 0; r < st.npaths; ++r) acc = t16(acc + int(st.paths[r * vol + i]));
            S[i] = static_cast<uint16_t>(acc);
        }
    });
    std::vector<int16_t> dl(n), dr(n);
    int16_t* wi = dumps ? dumps->wta_int : nullptr;
    parallel_bands(n, st.threads, [&](size_t p0, size_t p1) {
        wta_left_pixels(S, st.D, st.m, st.p->uniquenessRatio, dl.data(), wi, p0, p1);
    });
    parallel_bands(st.H, st.threads, [&](size_t y0, size_t y1) {
        wta_right_rows(S, st.W, st.D, st.m, dr.data(), static_cast<int>(y0), static_cast<int>(y1));
    });
    if (dumps) {
        if (dumps->disp_left) std::memcpy(dumps->disp_left, dl.data(), n * 2);
        if (dumps->disp_right) std::memcpy(dumps->disp_right, dr.data(), n * 2);
    }
    if (st.p->disp12MaxDiff >= 0)
        oracle_lr_check(dl.data(), dr.data(), st.W, st.H, st.p->disp12MaxDiff * 16);
    if (dumps && dumps->after_lr) std::memcpy(dumps->after_lr, dl.data(), n * 2);
    if (st.p->speckleWindowSize > 0)
        oracle_speckle(dl.data(), st.W, st.H, std::min(st.p->speckleWindowSize, 25),
                       st.p->speckleRange * 16);
    if (dumps && dumps->after_speckle) std::memcpy(dumps->after_speckle, dl.data(), n * 2);
    oracle_median(dl.data(), out, st.W, st.H);
}

void setup(Pipeline& st, const oracle_params* p, int W, int H, int threads,
           const oracle_dumps* dumps) {
    st.p = p;
    st.W = W;
    st.H = H;
    st.m = p->minDisparity;
    st.D = p->maxDisparity - p->minDisparity;
    st.npaths = oracle_num_paths(p->numDirections);
    st.cs = oracle_cost_scale(p->matchingCostMode, p->blockSize, p->preFilterCap);
    st.threads = std::max(1, threads);
    const size_t vol = static_cast<size_t>(W) * H * st.D;
    st.cost = dumps && dumps->cost ? dumps->cost : (st.ownCost.resize(vol), st.ownCost.data());
    st.paths = dumps && dumps->paths ? dumps->paths
                                     : (st.ownPaths.resize(vol * st.npaths), st.ownPaths.data());
    st.sum = dumps && dumps->sum ? dumps->sum : (st.ownSum.resize(vol), st.ownSum.data());
}

int run(const oracle_params* p, const uint8_t* left, const uint8_t* right, int W, int H,
        int16_t* out, const oracle_dumps* dumps, int threads) {
    if (!valid_args(p, left, right, W, H) || !out) return 1;
    Pipeline st;
    setup(st, p, W, H, threads, dumps);
    cost_stage(st, left, right);
    path_stage(st, left);
    finish(st, out, dumps);
    return 0;
}

}  // namespace

extern "C" {

/* StereoSGBM::compute (src/sgbm_gpu.cu:147-195) for grayscale input. */
int oracle_sgbm(const oracle_params* p, const uint8_t* left, const uint8_t* right, int W, int H,
                int16_t* out, const oracle_dumps* dumps) {
    return run(p, left, right, W, H, out, dumps, 1);
}

int oracle_sgbm_mt(const oracle_params* p, const uint8_t* left, const uint8_t* right, int W,
                   int H, int16_t* out, int nthreads, const oracle_dumps* dumps) {
    return run(p, left, right, W, H, out, dumps, nthreads);
}

}  // extern "C"
