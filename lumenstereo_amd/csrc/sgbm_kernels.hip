// sgbm_kernels.hip - gfx950 kernels of the SGBM matcher.
//
// Layouts (all in HBM, row-major):
//   images            uint8  [H][W]
//   census            uint32 [H][W]
//   cost volume C     uint8  [H][W][Dp]   disparity innermost, Dp = D rounded up to 64
//   path costs L_r    uint16 [npaths][H][W][Dp]
//   disparity maps    int16  [H][W]       1/16 px, -1 invalid
// Padding entries d >= D are never read as costs: the path kernel forces them to 65535.
//
// Every integer operation reproduces the reference's arithmetic exactly, including the
// truncations to uint16_t (src/sgbm_kernels.cu:338-360); packed 16-bit VALU operations wrap
// mod 2^16 as those truncations do.
#include "sgbm_internal.h"

#include <hip/hip_runtime.h>

namespace lumen {

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t bits(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ u16x2 as16(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
__device__ __forceinline__ u16x2 splat(uint32_t v) { return as16((v & 0xFFFFu) * 0x10001u); }
__device__ __forceinline__ u16x2 vmin(u16x2 a, u16x2 b) { return __builtin_elementwise_min(a, b); }

// DPP controls (GFX9 encoding).
constexpr int kQuadXor1 = 0xB1;       // quad_perm [1,0,3,2]
constexpr int kQuadXor2 = 0x4E;       // quad_perm [2,3,0,1]
constexpr int kRowShl1 = 0x101;       // lane i reads lane i+1 of its row
constexpr int kRowShr1 = 0x111;       // lane i reads lane i-1 of its row
constexpr int kRowMirror = 0x140;     // lane i reads lane 15-i of its row
constexpr int kRowHalfMirror = 0x141; // lane i reads lane 7-i of its half row

template <int CTRL>
__device__ __forceinline__ uint32_t dpp(uint32_t v, uint32_t old) {
    return static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(static_cast<int>(old), static_cast<int>(v), CTRL, 0xF, 0xF, false));
}

// Minimum over the 16 lanes of a DPP row, left in every lane of the row.
__device__ __forceinline__ uint32_t row_min(uint32_t v) {
    v = min(v, dpp<kQuadXor1>(v, v));
    v = min(v, dpp<kQuadXor2>(v, v));
    v = min(v, dpp<kRowHalfMirror>(v, v));
    v = min(v, dpp<kRowMirror>(v, v));
    return v;
}

// Value of lane u of the lane's DPP row (row_newbcast, gfx90a and later); u is folded to a
// constant after unrolling.
__device__ __forceinline__ uint32_t row_bcast(uint32_t v, int u) {
    switch (u) {
        case 0: return dpp<0x150>(v, v);
        case 1: return dpp<0x151>(v, v);
        case 2: return dpp<0x152>(v, v);
        case 3: return dpp<0x153>(v, v);
        case 4: return dpp<0x154>(v, v);
        case 5: return dpp<0x155>(v, v);
        case 6: return dpp<0x156>(v, v);
        case 7: return dpp<0x157>(v, v);
        case 8: return dpp<0x158>(v, v);
        case 9: return dpp<0x159>(v, v);
        case 10: return dpp<0x15A>(v, v);
        case 11: return dpp<0x15B>(v, v);
        case 12: return dpp<0x15C>(v, v);
        case 13: return dpp<0x15D>(v, v);
        case 14: return dpp<0x15E>(v, v);
        default: return dpp<0x15F>(v, v);
    }
}

// float -> int16 as the reference's device conversion: truncate, clamp to the int16 range.
__device__ __forceinline__ int16_t f2i16(float v) {
    float t = truncf(v);
    t = fminf(fmaxf(t, -32768.0f), 32767.0f);
    return static_cast<int16_t>(static_cast<int>(t));
}

// Parabola fit of src/sgbm_kernels.cu:789-804 / :836-855 in IEEE single precision.
__device__ __forceinline__ float parabola(float s, uint32_t a0, uint32_t a1, uint32_t a2) {
    const float c0 = static_cast<float>(a0);
    const float c1 = static_cast<float>(a1);
    const float c2 = static_cast<float>(a2);
    const float denom = 2.0f * (c0 + c2 - 2.0f * c1);
    if (fabsf(denom) > 1e-6f) {
        float delta = (c0 - c2) / denom;
        delta = fmaxf(-1.0f, fminf(1.0f, delta));
        s += delta;
    }
    return s;
}

// ---------------------------------------------------------------------------------------
// Census cost (src/sgbm_kernels.cu:217-286), census transform fused in. One block per
// (row, 256-pixel segment): the five image rows around the row are staged in LDS, the 24-bit
// census codes of the segment and of every right pixel it can reach are computed into LDS
// (bit k over the 5x5 neighbours in row-major order, centre skipped, set when the neighbour
// is >= the centre; code 0 within 2 px of the border), then each thread writes 16
// consecutive disparities of one pixel as one 16-byte store.
// ---------------------------------------------------------------------------------------
constexpr int kCensusTile = 256;  // pixels per block of the census cost kernel
constexpr int kCostTile = 64;     // pixels per block of the SAD cost kernel

__device__ __forceinline__ uint32_t census_at(const uint8_t* rows, int pitch, int c) {
    // rows: 5 staged image rows, column c is the centre's column in the staging buffer.
    const uint8_t ctr = rows[2 * pitch + c];
    uint32_t code = 0;
    int k = 0;
#pragma unroll
    for (int dy = 0; dy < 5; ++dy)
#pragma unroll
        for (int dx = -2; dx <= 2; ++dx) {
            if (dy == 2 && dx == 0) continue;
            code |= (rows[dy * pitch + c + dx] >= ctr ? 1u : 0u) << k;
            ++k;
        }
    return code;
}

__global__ __launch_bounds__(256) void cost_census_kernel(const uint8_t* __restrict__ left,
                                                          const uint8_t* __restrict__ right,
                                                          uint8_t* __restrict__ cost, int W,
                                                          int H, int D, int Dp, int m) {
    constexpr int T = kCensusTile;
    extern __shared__ uint32_t shc[];
    const int y = blockIdx.y, x0 = blockIdx.x * T;
    const int rs = x0 - m - (Dp - 1);           // first right pixel the segment can reach
    const int nr = T + Dp - 1;                  // right pixels the segment can reach
    const int lpitch = T + 4, rpitch = nr + 4;
    uint32_t* sl = shc;                         // T codes
    uint32_t* sr = shc + T;                     // nr codes
    uint8_t* il = reinterpret_cast<uint8_t*>(shc + T + nr);  // 5 x lpitch bytes
    uint8_t* ir = il + 5 * lpitch;                             // 5 x rpitch bytes
    const bool rowOk = y >= 2 && y < H - 2;
    if (rowOk) {
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            const uint8_t* lrow = left + (size_t)(y - 2 + r) * W;
            const uint8_t* rrow = right + (size_t)(y - 2 + r) * W;
            for (int i = threadIdx.x; i < lpitch; i += 256) {
                const int c = x0 - 2 + i;
                il[r * lpitch + i] = (c >= 0 && c < W) ? lrow[c] : uint8_t(0);
            }
            for (int i = threadIdx.x; i < rpitch; i += 256) {
                const int c = rs - 2 + i;
                ir[r * rpitch + i] = (c >= 0 && c < W) ? rrow[c] : uint8_t(0);
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < T; i += 256) {
            const int x = x0 + i;
            sl[i] = (x >= 2 && x < W - 2) ? census_at(il, lpitch, i + 2) : 0u;
        }
        for (int i = threadIdx.x; i < nr; i += 256) {
            const int rx = rs + i;
            sr[i] = (rx >= 2 && rx < W - 2) ? census_at(ir, rpitch, i + 2) : 0u;
        }
        __syncthreads();
    }
    const int chunks = Dp / 16;
    // item = pixel * chunks + chunk, advanced by 256 items without a division per item.
    int i = threadIdx.x / chunks, c = threadIdx.x - i * chunks;
    const int di = 256 / chunks, dc = 256 - di * chunks;
    for (; i < T; i += di, c += dc) {
        if (c >= chunks) {
            c -= chunks;
            ++i;
            if (i >= T) break;
        }
        const int x = x0 + i;
        if (x >= W) break;
        const bool leftOk = rowOk && x >= 2 && x < W - 2;
        const uint32_t a = leftOk ? sl[i] : 0u;
        const int d0 = c * 16;
        const uint32_t* rc = sr + i + Dp - 1 - d0;   // rc[-j] is the right code for d0 + j
        uint32_t w[4];
        if (leftOk && d0 + 15 < D && x - m - d0 - 15 >= 2 && x - m - d0 < W - 2) {
            // Interior: all 16 disparities valid. (h*255+12)/24 == (h*2718+168) >> 8 for
            // h in [0,24]; two results per packed 16-bit multiply-add, byte 1 of each half
            // gathered with one byte permute per four results.
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t h0 = __builtin_popcount(a ^ rc[-(4 * q + 0)]);
                const uint32_t h1 = __builtin_popcount(a ^ rc[-(4 * q + 1)]);
                const uint32_t h2 = __builtin_popcount(a ^ rc[-(4 * q + 2)]);
                const uint32_t h3 = __builtin_popcount(a ^ rc[-(4 * q + 3)]);
                const u16x2 k = splat(2718u), b = splat(168u);
                const u16x2 v01 = as16(h0 | (h1 << 16)) * k + b;
                const u16x2 v23 = as16(h2 | (h3 << 16)) * k + b;
                w[q] = __builtin_amdgcn_perm(bits(v23), bits(v01), 0x07050301u);
            }
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                uint32_t word = 0;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int d = d0 + q * 4 + b;
                    const int rx = x - m - d;
                    uint32_t v = 255;
                    if (leftOk && d < D && rx >= 2 && rx < W - 2) {
                        const uint32_t h = __builtin_popcount(a ^ rc[-(4 * q + b)]);
                        v = (h * 255u + 12u) / 24u;
                    }
                    word |= v << (8 * b);
                }
                w[q] = word;
            }
        }
        *reinterpret_cast<u32x4*>(cost + ((size_t)y * W + x) * Dp + d0) =
            u32x4{w[0], w[1], w[2], w[3]};
    }
}

// ---------------------------------------------------------------------------------------
// Sobel-x prefilter (src/sgbm_kernels.cu:38-59) fused into the SAD cost
// (src/sgbm_kernels.cu:172-212). One block per (row, 64-pixel segment); the prefiltered rows
// y-r..y+r of the left segment and of the right columns it can reach are built in LDS. The
// valid taps of a window form a rectangle, so count = rows * columns.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ uint8_t prefilter_at(const uint8_t* __restrict__ row, int c, int W,
                                                int cap) {
    int g;
    if (c > 0 && c < W - 1)
        g = int(row[c + 1]) - int(row[c - 1]);
    else if (c == 0)
        g = int(row[1]) - int(row[0]);
    else
        g = int(row[c]) - int(row[c - 1]);
    g = min(cap, max(-cap, g));
    return static_cast<uint8_t>(g + cap);
}

__global__ __launch_bounds__(256) void cost_sad_kernel(const uint8_t* __restrict__ left,
                                                       const uint8_t* __restrict__ right,
                                                       uint8_t* __restrict__ cost, int W, int H,
                                                       int D, int Dp, int m, int r, int cap) {
    extern __shared__ uint8_t shs[];
    const int lw = kCostTile + 2 * r;
    const int rw = kCostTile + 2 * r + Dp - 1;
    const int rows = 2 * r + 1;
    uint8_t* fl = shs;
    uint8_t* fr = shs + rows * lw;
    const int y = blockIdx.y, x0 = blockIdx.x * kCostTile;
    const int ls = x0 - r, rs = x0 - r - m - (Dp - 1);
    for (int i = threadIdx.x; i < rows * lw; i += 256) {
        const int k = i / lw, c = ls + (i - k * lw), yy = y - r + k;
        fl[i] = (c >= 0 && c < W && yy >= 0 && yy < H)
                    ? prefilter_at(left + (size_t)yy * W, c, W, cap) : uint8_t(0);
    }
    for (int i = threadIdx.x; i < rows * rw; i += 256) {
        const int k = i / rw, c = rs + (i - k * rw), yy = y - r + k;
        fr[i] = (c >= 0 && c < W && yy >= 0 && yy < H)
                    ? prefilter_at(right + (size_t)yy * W, c, W, cap) : uint8_t(0);
    }
    __syncthreads();
    const int dylo = max(-r, -y), dyhi = min(r, H - 1 - y);
    const int nrows = dyhi - dylo + 1;
    const int chunks = Dp / 16;
    for (int item = threadIdx.x; item < kCostTile * chunks; item += 256) {
        const int i = item / chunks, c = item - i * chunks;
        const int x = x0 + i;
        if (x >= W) continue;
        const int dxhi = min(r, W - 1 - x);
        const int d0 = c * 16;
        uint32_t w[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            uint32_t word = 0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int d = d0 + q * 4 + b;
                uint32_t v = 255;
                if (d < D) {
                    const int a = m + d;
                    const int dxlo = max(-r, a - x);
                    const int ncols = dxhi - dxlo + 1;
                    if (ncols > 0 && nrows > 0) {
                        int sum = 0;
                        for (int dy = dylo; dy <= dyhi; ++dy) {
                            const uint8_t* lrow = fl + (r + dy) * lw + i + r;
                            const uint8_t* rrow = fr + (r + dy) * rw + i + r + Dp - 1 - d;
                            for (int dx = dxlo; dx <= dxhi; ++dx)
                                sum += abs(int(lrow[dx]) - int(rrow[dx]));
                        }
                        v = static_cast<uint32_t>(sum / (ncols * nrows));
                    }
                }
                word |= (v & 0xFFu) << (8 * b);
            }
            w[q] = word;
        }
        *reinterpret_cast<u32x4*>(cost + ((size_t)y * W + x) * Dp + d0) =
            u32x4{w[0], w[1], w[2], w[3]};
    }
}

// Large windows (radius > 31) fall back to reading the images from global memory.
__global__ __launch_bounds__(256) void cost_sad_global_kernel(const uint8_t* __restrict__ left,
                                                              const uint8_t* __restrict__ right,
                                                              uint8_t* __restrict__ cost, int W,
                                                              int H, int D, int Dp, int m, int r,
                                                              int cap) {
    const size_t item = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t total = (size_t)W * H * Dp;
    if (item >= total) return;
    const int d = static_cast<int>(item % Dp);
    const size_t pix = item / Dp;
    const int x = static_cast<int>(pix % W), y = static_cast<int>(pix / W);
    uint32_t v = 255;
    if (d < D) {
        const int a = m + d;
        const int dylo = max(-r, -y), dyhi = min(r, H - 1 - y);
        const int dxlo = max(-r, a - x), dxhi = min(r, W - 1 - x);
        const int n = (dyhi - dylo + 1) * (dxhi - dxlo + 1);
        if (dyhi >= dylo && dxhi >= dxlo) {
            int sum = 0;
            for (int dy = dylo; dy <= dyhi; ++dy) {
                const uint8_t* lr = left + (size_t)(y + dy) * W;
                const uint8_t* rr = right + (size_t)(y + dy) * W;
                for (int dx = dxlo; dx <= dxhi; ++dx)
                    sum += abs(int(prefilter_at(lr, x + dx, W, cap)) -
                               int(prefilter_at(rr, x + dx - a, W, cap)));
            }
            v = static_cast<uint32_t>(sum / n);
        }
    }
    cost[item] = static_cast<uint8_t>(v);
}

// ---------------------------------------------------------------------------------------
// SGM aggregation: all paths of a frame in one launch.
//
// A scanline is the sequence of pixels of one path from its origin (the pixel whose
// predecessor lies outside the image) to the far border. Each scanline is owned by one DPP row
// (16 lanes); lane l holds disparities [l*K, l*K+K) as K/2 packed uint16 pairs, so the DP step
// is packed 16-bit VALU work, the d-1 / d+1 neighbours cross lanes with one row_shr/row_shl
// each, and min over d is a 4-step DPP butterfly. A wave advances 4 neighbouring scanlines of
// the same path in lockstep; scanline geometry is origin + t*step in flat pixel index.
// ---------------------------------------------------------------------------------------
__host__ __device__ inline int scan_count(int path, int W, int H) {
    return path < 2 ? H : (path < 4 ? W : W + H - 1);
}

int scanline_count(int path, int W, int H) { return scan_count(path, W, H); }

__device__ __forceinline__ void scan_geom(int path, int s, int W, int H, int& p0, int& step,
                                          int& len) {
    p0 = 0;
    step = 0;
    len = 0;
    if (s >= scan_count(path, W, H)) return;
    int x, y, dx, dy;
    if (path < 2) {  // LR, RL
        y = s;
        x = path == 0 ? 0 : W - 1;
        dx = path == 0 ? 1 : -1;
        dy = 0;
        len = W;
    } else if (path < 4) {  // TB, BT
        x = s;
        y = path == 2 ? 0 : H - 1;
        dx = 0;
        dy = path == 2 ? 1 : -1;
        len = H;
    } else {  // diagonals, laid out in the TLBR frame and mirrored
        int bx, by;
        if (s < H - 1) {
            bx = 0;
            by = H - 1 - s;
            len = min(W, s + 1);
        } else {
            bx = s - (H - 1);
            by = 0;
            len = min(W - bx, H);
        }
        const bool fx = (path == 5 || path == 6), fy = (path == 5 || path == 7);
        x = fx ? W - 1 - bx : bx;
        y = fy ? H - 1 - by : by;
        dx = fx ? -1 : 1;
        dy = fy ? -1 : 1;
    }
    p0 = y * W + x;
    step = dy * W + dx;
}

template <int K>
struct CostBytes {
    uint32_t w[K / 4];
};

template <int K>
__device__ __forceinline__ CostBytes<K> load_cost(const uint8_t* p) {
    CostBytes<K> c;
    if constexpr (K % 16 == 0) {
#pragma unroll
        for (int i = 0; i < K / 16; ++i) {
            const u32x4 v = reinterpret_cast<const u32x4*>(p)[i];
            c.w[4 * i + 0] = v.x;
            c.w[4 * i + 1] = v.y;
            c.w[4 * i + 2] = v.z;
            c.w[4 * i + 3] = v.w;
        }
    } else if constexpr (K % 8 == 0) {
#pragma unroll
        for (int i = 0; i < K / 8; ++i) {
            const u32x2 v = reinterpret_cast<const u32x2*>(p)[i];
            c.w[2 * i + 0] = v.x;
            c.w[2 * i + 1] = v.y;
        }
    } else {
#pragma unroll
        for (int i = 0; i < K / 4; ++i) c.w[i] = reinterpret_cast<const uint32_t*>(p)[i];
    }
    return c;
}

template <int K>
__device__ __forceinline__ void store_path(uint16_t* p, const u16x2 (&v)[K / 2]) {
    if constexpr (K % 8 == 0) {
#pragma unroll
        for (int i = 0; i < K / 8; ++i)
            __builtin_nontemporal_store(
                u32x4{bits(v[4 * i]), bits(v[4 * i + 1]), bits(v[4 * i + 2]), bits(v[4 * i + 3])},
                reinterpret_cast<u32x4*>(p) + i);
    } else {
#pragma unroll
        for (int i = 0; i < K / 4; ++i)
            __builtin_nontemporal_store(u32x2{bits(v[2 * i]), bits(v[2 * i + 1])},
                                        reinterpret_cast<u32x2*>(p) + i);
    }
}

// uint8 costs -> packed t16(C * costScale).
template <int K>
__device__ __forceinline__ void expand_cost(const CostBytes<K>& c, u16x2 cs, u16x2 (&out)[K / 2]) {
#pragma unroll
    for (int i = 0; i < K / 4; ++i) {
        out[2 * i] = as16(__builtin_amdgcn_perm(0u, c.w[i], 0x0c010c00u)) * cs;
        out[2 * i + 1] = as16(__builtin_amdgcn_perm(0u, c.w[i], 0x0c030c02u)) * cs;
    }
}

// d-1 and d+1 neighbour vectors of a packed per-lane slice (lane l holds d = l*K .. l*K+K-1).
template <int K>
__device__ __forceinline__ void neighbours(const u16x2 (&v)[K / 2], uint32_t (&dm1)[K / 2],
                                           uint32_t (&dp1)[K / 2]) {
    const uint32_t fromPrev = dpp<kRowShr1>(bits(v[K / 2 - 1]), 0xFFFFFFFFu);
    const uint32_t fromNext = dpp<kRowShl1>(bits(v[0]), 0xFFFFFFFFu);
#pragma unroll
    for (int j = 0; j < K / 2; ++j) {
        const uint32_t cur = bits(v[j]);
        const uint32_t lo = j == 0 ? fromPrev : bits(v[j - 1]);
        const uint32_t hi = j == K / 2 - 1 ? fromNext : bits(v[j + 1]);
        dm1[j] = (cur << 16) | (lo >> 16);
        dp1[j] = (cur >> 16) | (hi << 16);
    }
}

template <int K>
__device__ __forceinline__ uint32_t slice_min(const u16x2 (&v)[K / 2]) {
    u16x2 mm = v[0];
#pragma unroll
    for (int j = 1; j < K / 2; ++j) mm = vmin(mm, v[j]);
    const uint32_t b = bits(mm);
    return min(b & 0xFFFFu, b >> 16);
}

template <int K, bool PAD>
__global__ __launch_bounds__(256) void sgm_paths_kernel(PathsArgs a) {
    __shared__ uint16_t p2[256];
    p2[threadIdx.x] = a.p2tab[threadIdx.x];
    __syncthreads();
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wave >= a.waveStart[a.npaths]) return;
    int path = 0;
#pragma unroll
    for (int r = 1; r < kMaxPaths; ++r)
        if (r < a.npaths && wave >= a.waveStart[r]) path = r;
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, l = lane & 15;
    const int s = (wave - a.waveStart[path]) * 4 + g;
    int p0, step, len;
    scan_geom(path, s, a.W, a.H, p0, step, len);
    // The launch ends when its longest scanlines end, so waves with longer scanlines get
    // higher issue priority within their SIMD: horizontal > vertical > diagonal when W >= H.
    {
        const int ph = a.W >= a.H ? (path < 2 ? 3 : (path < 4 ? 2 : 1))
                                  : (path < 2 ? 2 : (path < 4 ? 3 : 1));
        if (ph == 3)
            __builtin_amdgcn_s_setprio(3);
        else if (ph == 2)
            __builtin_amdgcn_s_setprio(2);
        else
            __builtin_amdgcn_s_setprio(1);
    }
    int maxLen = max(len, __shfl_xor(len, 16));
    maxLen = max(maxLen, __shfl_xor(maxLen, 32));
    maxLen = __builtin_amdgcn_readfirstlane(maxLen);

    const int Dp = a.Dp, D = a.D;
    const uint8_t* cbase = a.cost + l * K;
    uint16_t* Lbase = a.L + (size_t)path * a.pathStride + l * K;
    const u16x2 cs = splat(static_cast<uint32_t>(a.costScale));
    const u16x2 P1p = splat(static_cast<uint32_t>(a.P1));
    const uint32_t m0 = l == 0 ? 0xFFFFu : 0u;  // slot d = 0 has no d-1 neighbour
    uint32_t c2m[K / 2], padm[K / 2];
#pragma unroll
    for (int j = 0; j < K / 2; ++j) {
        const int dlo = l * K + 2 * j;
        c2m[j] = (dlo == D - 1 ? 0xFFFFu : 0u) | (dlo + 1 == D - 1 ? 0xFFFF0000u : 0u);
        padm[j] = (dlo >= D ? 0xFFFFu : 0u) | (dlo + 1 >= D ? 0xFFFF0000u : 0u);
    }

    // Origin: L = t16(C * costScale).
    const long long cstep = (long long)step * Dp;  // bytes between consecutive cost vectors
    const uint8_t* lp = cbase + (size_t)p0 * Dp;   // cost pointer of the last issued load
    int lt = 0;                                     // step index of the last issued load
    uint16_t* sp = Lbase + (size_t)p0 * Dp;         // path-cost pointer of the current step
    uint16_t* const trash = a.trash + lane * K;     // stores of finished rows land here
    u16x2 Lv[K / 2];
    {
        const CostBytes<K> c0 = load_cost<K>(lp);
        expand_cost<K>(c0, cs, Lv);
        if constexpr (PAD) {
#pragma unroll
            for (int j = 0; j < K / 2; ++j) Lv[j] = as16(bits(Lv[j]) | padm[j]);
        }
        store_path<K>(len > 0 ? sp : trash, Lv);
    }
    uint32_t M = row_min(slice_min<K>(Lv));

    // Two register blocks of PF steps alternate: while one block is computed, the costs of the
    // next are in flight. Lane l of a row also fetches the two image bytes of step
    // tb + (l % PF) and looks up t16(P2adj) once per block; each step broadcasts its lane's
    // value with DPP row_newbcast. Loads past the end of a scanline re-read its last pixel;
    // steps past the end are computed but stored to the trash slot, so every store is
    // unconditional and the wait counts stay exact.
    constexpr int PF = K <= 8 ? 8 : (K <= 16 ? 4 : 2);
    const int last = len > 0 ? len - 1 : 0;
    const int lu = l & (PF - 1);
    auto fetch = [&](int tb, CostBytes<K>(&dst)[PF], int (&iv)[2]) {
        const int ti = min(tb + lu, last);
        iv[0] = a.img[p0 + ti * step];
        iv[1] = a.img[p0 + max(ti - 1, 0) * step];
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const bool adv = lt < last;
            lp = adv ? lp + cstep : lp;
            lt += adv ? 1 : 0;
            dst[u] = load_cost<K>(lp);
        }
    };
    auto run = [&](int tb, const CostBytes<K>(&blk)[PF], const int (&iv)[2]) {
        const uint32_t p2v = p2[abs(iv[0] - iv[1])];
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int t = tb + u;
            const bool act = t < len;
            sp += step * Dp;
            const u16x2 c3 = splat(M + row_bcast(p2v, u));
            const u16x2 Mp = splat(M);
            u16x2 Ch[K / 2];
            expand_cost<K>(blk[u], cs, Ch);
            uint32_t dm1[K / 2], dp1[K / 2];
            neighbours<K>(Lv, dm1, dp1);
#pragma unroll
            for (int j = 0; j < K / 2; ++j) {
                uint32_t c1 = bits(as16(dm1[j]) + P1p);
                uint32_t c2 = bits(as16(dp1[j]) + P1p);
                if (j == 0) c1 |= m0;
                if constexpr (PAD) {
                    c2 |= c2m[j];
                } else {
                    if (j == K / 2 - 1) c2 |= c2m[j];
                }
                const u16x2 mn = vmin(vmin(Lv[j], as16(c1)), vmin(as16(c2), c3));
                u16x2 nv = Ch[j] + mn - Mp;
                if constexpr (PAD) nv = as16(bits(nv) | padm[j]);
                Lv[j] = nv;
            }
            store_path<K>(act ? sp : trash, Lv);
            M = row_min(slice_min<K>(Lv));
        }
    };
    CostBytes<K> bufA[PF], bufB[PF];
    int gA[2], gB[2];
    fetch(1, bufA, gA);
    // Entering the loop with block A complete keeps the loop-header state identical on both
    // edges, so no full drain is needed there.
    __builtin_amdgcn_s_waitcnt((0 & 0xF) | (7 << 4) | (15 << 8));
    for (int t0 = 1; t0 < maxLen; t0 += 2 * PF) {
        // The scheduling barriers keep each block's loads where they are issued: moved next to
        // their first use they would stall there.
        fetch(t0 + PF, bufB, gB);
        __builtin_amdgcn_sched_barrier(0);
        run(t0, bufA, gA);
        fetch(t0 + 2 * PF, bufA, gA);
        __builtin_amdgcn_sched_barrier(0);
        run(t0 + PF, bufB, gB);
        // Block A's loads were issued before block B's stores: waiting until only those
        // stores may be outstanding retires the loads without waiting on the stores.
        // (s_waitcnt vmcnt(N) expcnt(7) lgkmcnt(15), GFX9 encoding.)
        constexpr int N = PF * (K % 8 == 0 ? K / 8 : K / 4);
        static_assert(N < 64, "vmcnt is 6 bits");
        __builtin_amdgcn_s_waitcnt((N & 0xF) | (7 << 4) | (15 << 8) | (((N >> 4) & 3) << 14));
    }
}

template <int K, bool PAD>
static hipError_t launch_paths_k(PathsArgs& a, int blocks, hipStream_t s) {
    hipLaunchKernelGGL((sgm_paths_kernel<K, PAD>), dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

template <int K>
static hipError_t launch_paths_pad(PathsArgs& a, int blocks, hipStream_t s) {
    return a.D == a.Dp ? launch_paths_k<K, false>(a, blocks, s) : launch_paths_k<K, true>(a, blocks, s);
}

hipError_t launch_paths(PathsArgs& a, hipStream_t s) {
    int waves = 0;
    for (int r = 0; r < a.npaths; ++r) {
        a.waveStart[r] = waves;
        waves += (scan_count(r, a.W, a.H) + 3) / 4;
    }
    for (int r = a.npaths; r <= kMaxPaths; ++r) a.waveStart[r] = waves;
    const int blocks = (waves + 3) / 4;
    switch (a.Dp / kScanLanes) {
        case 4: return launch_paths_pad<4>(a, blocks, s);
        case 8: return launch_paths_pad<8>(a, blocks, s);
        case 12: return launch_paths_pad<12>(a, blocks, s);
        case 16: return launch_paths_pad<16>(a, blocks, s);
        case 20: return launch_paths_pad<20>(a, blocks, s);
        case 24: return launch_paths_pad<24>(a, blocks, s);
        case 28: return launch_paths_pad<28>(a, blocks, s);
        case 32: return launch_paths_pad<32>(a, blocks, s);
        default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------------------
// Path sum + left WTA + right WTA + LR check, one block per image row.
//
// Each DPP row of the block takes one pixel at a time: it sums the path costs (mod 2^16),
// finds the left winner with a 4-step butterfly over (cost<<16 | d) keys plus the second
// minimum (src/sgbm_kernels.cu:748-805), and offers every (pixel, d) to right pixel
// x - m - d as the 64-bit key cost<<48 | d<<32 | S(d-1)<<16 | S(d+1) through an LDS atomic
// min. Smallest key = smallest cost, then smallest d, which is the first-minimum rule of the
// right WTA (src/sgbm_kernels.cu:820-834); costs equal to 65535 are never offered because
// they never beat the reference's initial value. The LR check (src/sgbm_kernels.cu:862-890)
// only reads the same row, so it runs here too.
// ---------------------------------------------------------------------------------------
template <int K>
__device__ __forceinline__ uint32_t pick(const u16x2 (&v)[K / 2], int slot) {
    uint32_t r = 0xFFFFu;
#pragma unroll
    for (int j = 0; j < K / 2; ++j) {
        const uint32_t b = bits(v[j]);
        r = slot == 2 * j ? (b & 0xFFFFu) : r;
        r = slot == 2 * j + 1 ? (b >> 16) : r;
    }
    return r;
}

template <int K>
__device__ __forceinline__ uint32_t pick32(const uint32_t (&v)[K / 2], int slot) {
    uint32_t r = 0xFFFFu;
#pragma unroll
    for (int j = 0; j < K / 2; ++j) {
        r = slot == 2 * j ? (v[j] & 0xFFFFu) : r;
        r = slot == 2 * j + 1 ? (v[j] >> 16) : r;
    }
    return r;
}

constexpr int kWtaThreads = 256;

template <int K, bool PAD>
__global__ __launch_bounds__(kWtaThreads) void wta_row_kernel(WtaArgs a) {
    // rbest[slot(xr)]: running right-WTA key of right pixel xr. One padding entry every K
    // entries spreads the 16 lanes of a row (whose candidates are K right pixels apart) over
    // distinct LDS banks.
    extern __shared__ unsigned long long rbest[];
    const int W = a.W, D = a.D, Dp = a.Dp, m = a.m;
    const int nslots = W + W / K + 1;
    int16_t* sdl = reinterpret_cast<int16_t*>(rbest + nslots);
    int16_t* sdr = sdl + W;
    const int y = blockIdx.x;
    for (int i = threadIdx.x; i < nslots; i += kWtaThreads) rbest[i] = ~0ull;
    __syncthreads();
    const int g = threadIdx.x >> 4, l = threadIdx.x & 15;
    uint32_t padm[K / 2];
#pragma unroll
    for (int j = 0; j < K / 2; ++j) {
        const int dlo = l * K + 2 * j;
        padm[j] = (dlo >= D ? 0xFFFFu : 0u) | (dlo + 1 >= D ? 0xFFFF0000u : 0u);
    }
    const size_t rowPix = (size_t)y * W;
    for (int x = g; x < W; x += kWtaThreads / 16) {
        const uint16_t* src = a.L + (rowPix + x) * Dp + l * K;
        u16x2 S[K / 2];
#pragma unroll
        for (int j = 0; j < K / 2; ++j) S[j] = as16(0u);
        for (int r = 0; r < a.npaths; ++r) {
            const uint16_t* q = src + (size_t)r * a.pathStride;
            if constexpr (K % 8 == 0) {
#pragma unroll
                for (int i = 0; i < K / 8; ++i) {
                    const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(q) + i);
                    S[4 * i + 0] += as16(v.x);
                    S[4 * i + 1] += as16(v.y);
                    S[4 * i + 2] += as16(v.z);
                    S[4 * i + 3] += as16(v.w);
                }
            } else {
#pragma unroll
                for (int i = 0; i < K / 4; ++i) {
                    const u32x2 v = __builtin_nontemporal_load(reinterpret_cast<const u32x2*>(q) + i);
                    S[2 * i + 0] += as16(v.x);
                    S[2 * i + 1] += as16(v.y);
                }
            }
        }
        if constexpr (PAD) {
#pragma unroll
            for (int j = 0; j < K / 2; ++j) S[j] = as16(bits(S[j]) | padm[j]);
        }
        uint32_t sm1[K / 2], sp1[K / 2];
        neighbours<K>(S, sm1, sp1);

        // Left WTA: (first minimum, second minimum of the multiset).
        uint32_t k1 = 0xFFFFFFFFu, m2 = 0xFFFFu;
#pragma unroll
        for (int j = 0; j < K / 2; ++j) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint32_t v = h ? (bits(S[j]) >> 16) : (bits(S[j]) & 0xFFFFu);
                const uint32_t key = (v << 16) | static_cast<uint32_t>(l * K + 2 * j + h);
                m2 = min(m2, max(k1, key) >> 16);
                k1 = min(k1, key);
            }
        }
#define LUMEN_WTA_COMBINE(CTRL)                                  \
    {                                                            \
        const uint32_t ok = dpp<CTRL>(k1, k1);                   \
        const uint32_t om = dpp<CTRL>(m2, m2);                   \
        m2 = min(min(m2, om), max(k1, ok) >> 16);                \
        k1 = min(k1, ok);                                        \
    }
        LUMEN_WTA_COMBINE(kQuadXor1)
        LUMEN_WTA_COMBINE(kQuadXor2)
        LUMEN_WTA_COMBINE(kRowHalfMirror)
        LUMEN_WTA_COMBINE(kRowMirror)
#undef LUMEN_WTA_COMBINE
        const int bestD = static_cast<int>(k1 & 0xFFFFu);
        const uint32_t best = k1 >> 16;
        if (l == bestD / K) {
            int16_t out;
            const uint32_t ratio = static_cast<uint32_t>(a.uniquenessRatio);
            if (a.uniquenessRatio > 0 && m2 != 0xFFFFu && m2 <= best + (best * ratio) / 100u) {
                out = -1;
            } else {
                float sub = static_cast<float>(bestD);
                if (bestD > 0 && bestD < D - 1) {
                    const int slot = bestD - l * K;
                    sub = parabola(sub, pick32<K>(sm1, slot), best, pick32<K>(sp1, slot));
                }
                out = f2i16((sub + static_cast<float>(m)) * 16.0f);
            }
            a.dispL[rowPix + x] = out;
            a.wtaInt[rowPix + x] = static_cast<int16_t>(bestD);
            sdl[x] = out;
        }

        // Right WTA candidates.
#pragma unroll
        for (int j = 0; j < K / 2; ++j) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int d = l * K + 2 * j + h;
                const int xr = x - m - d;
                const uint32_t v = h ? (bits(S[j]) >> 16) : (bits(S[j]) & 0xFFFFu);
                if (d < D && xr >= 0 && v != 0xFFFFu) {
                    const uint32_t lo = h ? (sm1[j] >> 16) : (sm1[j] & 0xFFFFu);
                    const uint32_t hi = h ? (sp1[j] >> 16) : (sp1[j] & 0xFFFFu);
                    const unsigned long long key = (static_cast<unsigned long long>(v) << 48) |
                                                   (static_cast<unsigned long long>(d) << 32) |
                                                   (lo << 16) | hi;
                    atomicMin(&rbest[xr + xr / K], key);
                }
            }
        }
    }
    __syncthreads();
    for (int xr = threadIdx.x; xr < W; xr += kWtaThreads) {
        const unsigned long long key = rbest[xr + xr / K];
        float sub = 0.0f;
        if (key != ~0ull) {
            const int bd = static_cast<int>((key >> 32) & 0xFFFFu);
            sub = static_cast<float>(bd);
            if (bd > 0 && bd < D - 1)
                sub = parabola(sub, static_cast<uint32_t>((key >> 16) & 0xFFFFu),
                               static_cast<uint32_t>(key >> 48), static_cast<uint32_t>(key & 0xFFFFu));
        }
        const int16_t dr = f2i16((sub + static_cast<float>(m)) * 16.0f);
        a.dispR[rowPix + xr] = dr;
        sdr[xr] = dr;
    }
    __syncthreads();
    for (int x = threadIdx.x; x < W; x += kWtaThreads) {
        int16_t dl = sdl[x];
        if (a.doLR && dl > 0) {
            const int xr = x - dl / 16;
            if (xr < 0 || xr >= W) {
                dl = -1;
            } else {
                const int16_t dr = sdr[xr];
                if (dr <= 0 || abs(int(dl) - int(dr)) > a.lrMaxDiff) dl = -1;
            }
        }
        a.dispLR[rowPix + x] = dl;
    }
}

template <int K, bool PAD>
static hipError_t launch_wta_k(const WtaArgs& a, hipStream_t s) {
    const size_t lds = (size_t)(a.W + a.W / K + 1) * 8 + (size_t)a.W * 4;
    hipLaunchKernelGGL((wta_row_kernel<K, PAD>), dim3(a.H), dim3(kWtaThreads), lds, s, a);
    return hipGetLastError();
}

template <int K>
static hipError_t launch_wta_pad(const WtaArgs& a, hipStream_t s) {
    return a.D == a.Dp ? launch_wta_k<K, false>(a, s) : launch_wta_k<K, true>(a, s);
}

hipError_t launch_wta(const WtaArgs& a, hipStream_t s) {
    if (a.W > kMaxWidth) return hipErrorInvalidValue;
    switch (a.Dp / kScanLanes) {
        case 4: return launch_wta_pad<4>(a, s);
        case 8: return launch_wta_pad<8>(a, s);
        case 12: return launch_wta_pad<12>(a, s);
        case 16: return launch_wta_pad<16>(a, s);
        case 20: return launch_wta_pad<20>(a, s);
        case 24: return launch_wta_pad<24>(a, s);
        case 28: return launch_wta_pad<28>(a, s);
        case 32: return launch_wta_pad<32>(a, s);
        default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------------------
// Speckle filter (src/sgbm_kernels.cu:979-1013) with snapshot semantics: reads `in`, writes
// `out`, so the result does not depend on thread timing.
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void speckle_kernel(const int16_t* __restrict__ in,
                                                      int16_t* __restrict__ out, int W, int H,
                                                      int minSimilar, int maxDiff) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    const int16_t c = in[(size_t)y * W + x];
    int16_t res = c;
    if (c > 0) {
        int n = 0;
        for (int dy = -2; dy <= 2; ++dy)
            for (int dx = -2; dx <= 2; ++dx) {
                const int nx = x + dx, ny = y + dy;
                if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
                const int16_t v = in[(size_t)ny * W + nx];
                if (v > 0 && abs(int(c) - int(v)) <= maxDiff) ++n;
            }
        if (n < minSimilar) res = -1;
    }
    out[(size_t)y * W + x] = res;
}

// ---------------------------------------------------------------------------------------
// 3x3 median of positive values (src/sgbm_kernels.cu:892-935): 64x4 tiles with a 1-pixel halo
// in LDS; out-of-image taps are staged as 0, which the positive-only rule excludes. The nine
// taps (non-positive ones replaced by INT_MAX so they sort last) go through a fixed
// 25-comparator sorting network; the upper median is element n/2 of the sorted positives.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ void cswap(int& a, int& b) {
    const int lo = min(a, b);
    b = max(a, b);
    a = lo;
}

__global__ __launch_bounds__(256) void median3x3_kernel(const int16_t* __restrict__ in,
                                                        int16_t* __restrict__ out, int W, int H) {
    constexpr int TW = 64, TH = 4;
    __shared__ int16_t tile[TH + 2][TW + 2];
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;
    for (int i = threadIdx.x; i < (TH + 2) * (TW + 2); i += 256) {
        const int ty = i / (TW + 2), tx = i % (TW + 2);
        const int gx = x0 + tx - 1, gy = y0 + ty - 1;
        tile[ty][tx] = (gx >= 0 && gx < W && gy >= 0 && gy < H) ? in[(size_t)gy * W + gx] : int16_t(0);
    }
    __syncthreads();
    const int tx = threadIdx.x % TW, ty = threadIdx.x / TW;
    const int x = x0 + tx, y = y0 + ty;
    if (x >= W || y >= H) return;
    int v[9];
    int n = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int t = tile[ty + k / 3][tx + k % 3];
        n += t > 0 ? 1 : 0;
        v[k] = t > 0 ? t : 0x7FFFFFFF;
    }
    cswap(v[0], v[1]); cswap(v[3], v[4]); cswap(v[6], v[7]);
    cswap(v[1], v[2]); cswap(v[4], v[5]); cswap(v[7], v[8]);
    cswap(v[0], v[1]); cswap(v[3], v[4]); cswap(v[6], v[7]);
    cswap(v[0], v[3]); cswap(v[3], v[6]); cswap(v[0], v[3]);
    cswap(v[1], v[4]); cswap(v[4], v[7]); cswap(v[1], v[4]);
    cswap(v[2], v[5]); cswap(v[5], v[8]); cswap(v[2], v[5]);
    cswap(v[1], v[3]); cswap(v[5], v[7]); cswap(v[2], v[6]);
    cswap(v[4], v[6]); cswap(v[2], v[4]); cswap(v[2], v[3]);
    cswap(v[5], v[6]);
    const int k = n >> 1;
    int r = v[0];
#pragma unroll
    for (int j = 1; j < 9; ++j) r = k == j ? v[j] : r;
    out[(size_t)y * W + x] = n == 0 ? tile[ty + 1][tx + 1] : static_cast<int16_t>(r);
}

// Sum of all path costs mod 2^16 into a padded volume (parity dumps only).
__global__ __launch_bounds__(256) void path_sum_kernel(const uint16_t* __restrict__ L,
                                                       size_t pathStride, int npaths,
                                                       uint16_t* __restrict__ S, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint32_t acc = 0;
    for (int r = 0; r < npaths; ++r) acc += L[(size_t)r * pathStride + i];
    S[i] = static_cast<uint16_t>(acc & 0xFFFFu);
}

// ---------------------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------------------
hipError_t launch_cost_census(const uint8_t* left, const uint8_t* right, uint8_t* cost, int W,
                              int H, int D, int Dp, int m, hipStream_t s) {
    dim3 grid((W + kCensusTile - 1) / kCensusTile, H);
    const int nr = kCensusTile + Dp - 1;
    const size_t lds = (size_t)(kCensusTile + nr) * 4 + 5 * (kCensusTile + 4) + 5 * (nr + 4);
    hipLaunchKernelGGL(cost_census_kernel, grid, dim3(256), lds, s, left, right, cost, W, H, D,
                       Dp, m);
    return hipGetLastError();
}

hipError_t launch_cost_sad(const uint8_t* left, const uint8_t* right, uint8_t* cost, int W, int H,
                           int D, int Dp, int m, int blockSize, int preFilterCap, hipStream_t s) {
    const int r = blockSize / 2;
    const int cap = preFilterCap > 1 ? preFilterCap : 1;
    if (r < 0) {
        // A negative radius gives an empty window: every cost is 255.
        return hipMemsetAsync(cost, 0xFF, (size_t)W * H * Dp, s);
    }
    if (r <= 31) {
        dim3 grid((W + kCostTile - 1) / kCostTile, H);
        const size_t lds = (size_t)(2 * r + 1) * ((kCostTile + 2 * r) + (kCostTile + 2 * r + Dp - 1));
        hipLaunchKernelGGL(cost_sad_kernel, grid, dim3(256), lds, s, left, right, cost, W, H, D, Dp,
                           m, r, cap);
    } else {
        const size_t total = (size_t)W * H * Dp;
        hipLaunchKernelGGL(cost_sad_global_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256),
                           0, s, left, right, cost, W, H, D, Dp, m, r, cap);
    }
    return hipGetLastError();
}

hipError_t launch_speckle(const int16_t* in, int16_t* out, int W, int H, int minSimilar,
                          int maxDiff, hipStream_t s) {
    dim3 grid((W + 15) / 16, (H + 15) / 16);
    hipLaunchKernelGGL(speckle_kernel, grid, dim3(256), 0, s, in, out, W, H, minSimilar, maxDiff);
    return hipGetLastError();
}

hipError_t launch_median(const int16_t* in, int16_t* out, int W, int H, hipStream_t s) {
    dim3 grid((W + 63) / 64, (H + 3) / 4);
    hipLaunchKernelGGL(median3x3_kernel, grid, dim3(256), 0, s, in, out, W, H);
    return hipGetLastError();
}

hipError_t launch_path_sum(const uint16_t* L, size_t pathStride, int npaths, uint16_t* S,
                           size_t n, hipStream_t s) {
    hipLaunchKernelGGL(path_sum_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, L,
                       pathStride, npaths, S, n);
    return hipGetLastError();
}

}  // namespace lumen
