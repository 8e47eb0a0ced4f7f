// sgbm_device.h - device code of the matcher, shared by the per-pair kernels
// (sgbm_kernels.hip) and the persistent batched kernels (sgbm_batch.hip).
//
// Every stage is a __device__ body over one unit of work of one pair (a row segment for the
// cost, a wave of four scanlines for the aggregation, a row for the WTA, a tile for the
// post-processing); the per-pair and batched kernels differ only in how they hand out those
// units, so their results are identical by construction.
//
// Layouts (all in HBM, row-major):
//   images            uint8  [H][W]
//   cost volume C     uint8  [H][W][Dp]   disparity innermost, Dp = D rounded up to 64
//   path costs L_r    uint16 [npaths][H][W][Dp]
//   disparity maps    int16  [H][W]       1/16 px, -1 invalid
// Padding entries d >= D carry cost 255 and are forced to 65535 in every path cost.
//
// Every integer operation reproduces the reference's arithmetic exactly, including the
// truncations to uint16_t (src/sgbm_kernels.cu:338-360); packed 16-bit VALU operations wrap
// mod 2^16 as those truncations do.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <type_traits>

#include "sgbm_internal.h"

namespace lumen {

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
// Vectors of dwords at 4-byte alignment (plain vector types assume their size, rounded up to
// a power of two).
typedef unsigned int u32x3 __attribute__((ext_vector_type(3), aligned(4)));
typedef unsigned int u32x2a __attribute__((ext_vector_type(2), aligned(4)));
typedef short i16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t bits(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ u16x2 as16(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
__device__ __forceinline__ u16x2 splat(uint32_t v) { return as16((v & 0xFFFFu) * 0x10001u); }
// The low half of v in both halves, as a shuffle: a packed 16-bit consumer reads it through
// its operand select (op_sel_hi), so no instruction builds the pair.
__device__ __forceinline__ u16x2 bcast_lo(uint32_t v) {
    const u16x2 t = as16(v);
    return __builtin_shufflevector(t, t, 0, 0);
}
__device__ __forceinline__ u16x2 vmin(u16x2 a, u16x2 b) { return __builtin_elementwise_min(a, b); }

// f(std::integral_constant<int, I>{}) for I = B .. E-1 in order: unrolled steps whose index is
// a compile-time constant (register ring slots, immediate LDS offsets).
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Global-memory view of a pointer. The stage bodies dereference the PairJob's buffers through
// it: a PairJob read from device memory (the batched kernels' pair table) carries generic
// pointers, for which the compiler emits flat accesses, and flat accesses also count against
// the LDS wait counter, so every LDS wait of a stage would wait for them as well.
#define LUMEN_G __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ LUMEN_G T* gmem(T* p) {
    return (LUMEN_G T*)p;
}

// DPP controls (GFX9 encoding). Each names a source lane of the same row for every lane.
constexpr int kQuadXor1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kQuadXor2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kRowRor1 = 0x121;        // lane i reads lane (i-1) mod 16 of its row
constexpr int kRowRor15 = 0x12F;       // lane i reads lane (i+1) mod 16 of its row
constexpr int kRowMirror = 0x140;      // lane i reads lane 15-i of its row
constexpr int kRowHalfMirror = 0x141;  // lane i reads lane 7-i of its half row
constexpr int kWaveRor1 = 0x13C;       // lane i reads lane (i-1) mod 64 of the wave
constexpr int kWaveRol1 = 0x134;       // lane i reads lane (i+1) mod 64 of the wave

// v of the lane CTRL names. Only controls that give every lane a source are used (the ones
// above and row_newbcast), so no lane keeps an old value and the bound_ctrl form is exact; it
// also lets the compiler fold the move into the VALU instruction that consumes it (v_min_u32
// with a DPP source), one instruction per butterfly stage instead of three.
//
// Contract: every caller keeps whole 16-lane rows active (no divergent branch around a DPP
// use that disables part of a row). A DPP read of a disabled source lane returns 0 in the
// bound_ctrl form instead of that lane's value, which would silently change a minimum. The
// aggregation meets this by running finished or missing scanlines on clamped inputs with their
// stores redirected, never by branching lanes off; the WTA hands each pixel to a whole row, so
// its loop conditions are uniform within a row.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
    return static_cast<uint32_t>(
        __builtin_amdgcn_mov_dpp(static_cast<int>(v), CTRL, 0xF, 0xF, true));
}

// Minimum over the 16 lanes of a DPP row, left in every lane of the row.
__device__ __forceinline__ uint32_t row_min(uint32_t v) {
    v = min(v, dpp<kQuadXor1>(v));
    v = min(v, dpp<kQuadXor2>(v));
    v = min(v, dpp<kRowHalfMirror>(v));
    v = min(v, dpp<kRowMirror>(v));
    return v;
}

// Value of lane u of the lane's DPP row (row_newbcast, gfx90a and later); u is folded to a
// constant after unrolling.
__device__ __forceinline__ uint32_t row_bcast(uint32_t v, int u) {
    switch (u) {
        case 0: return dpp<0x150>(v);
        case 1: return dpp<0x151>(v);
        case 2: return dpp<0x152>(v);
        case 3: return dpp<0x153>(v);
        case 4: return dpp<0x154>(v);
        case 5: return dpp<0x155>(v);
        case 6: return dpp<0x156>(v);
        case 7: return dpp<0x157>(v);
        case 8: return dpp<0x158>(v);
        case 9: return dpp<0x159>(v);
        case 10: return dpp<0x15A>(v);
        case 11: return dpp<0x15B>(v);
        case 12: return dpp<0x15C>(v);
        case 13: return dpp<0x15D>(v);
        case 14: return dpp<0x15E>(v);
        default: return dpp<0x15F>(v);
    }
}

// s_waitcnt vmcnt(N) with expcnt and lgkmcnt left open (GFX9 encoding).
#define LUMEN_WAIT_VMCNT(N) \
    __builtin_amdgcn_s_waitcnt(((N) & 0xF) | (7 << 4) | (15 << 8) | ((((N) >> 4) & 3) << 14))

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
// Census cost (src/sgbm_kernels.cu:217-286), census transform fused in: one row segment of
// CT pixels per call, block-wide. The five image rows around the row are staged in LDS, the
// 24-bit census codes of the segment and of every right pixel it can reach are computed into
// LDS (bit k over the 5x5 neighbours in row-major order, centre skipped, set when the
// neighbour is >= the centre; code 0 within 2 px of the border), then each thread writes 16
// consecutive disparities of one pixel as one 16-byte store.
// ---------------------------------------------------------------------------------------
constexpr int kCensusTile = 256;  // default pixels per census cost segment (batched form)
// Pixels per SAD cost segment. Wider segments (128, 256) were measured slower at every range
// tried (64 to 192): fewer blocks overlap less of each other's row staging.
constexpr int kSadTile = 64;

// Census output is computed 64 pixels (one per lane) at a time into an LDS tile whose rows are
// Dp + 16 bytes apart, so the 16-byte writes of 16 consecutive lanes start on 16 different
// 4-bank groups.
constexpr int kCensusPass = 64;

// LDS of a census segment of T pixels: the output tile of one pass, left codes, the codes of
// every right pixel the segment can reach, and the five staged rows of both images.
__host__ __device__ inline size_t census_lds_bytes(int Dp, int T = kCensusTile) {
    const int nr = T + Dp - 1;
    return (size_t)kCensusPass * (Dp + 16) + (size_t)(T + nr) * 4 + 5 * (T + 4) + 5 * (nr + 4);
}

// Windows of radius 1 to 3 (block sizes 3, 5, 7) run a different segment: every thread owns
// four consecutive disparities and slides over kSadRun pixels, and a 256-thread block covers
// 1024 / Dp such runs of one row.
constexpr int kSadRun = 16;
__host__ __device__ constexpr bool sad_runs_radius(int r) { return r >= 1 && r <= 3; }
__host__ __device__ inline int sad_segment_px(int r, int Dp) {
    return sad_runs_radius(r) ? kSadRun * (1024 / Dp) : kSadTile;
}

// (mul, add) of a window of n taps (0 <= n <= 49) such that floor(s / n) is byte 3 of
// s * mul + add (mul < 2^24, for a 24-bit multiply) for every s <= 255 n, and 255 for n = 0:
// - n >= 2: mul = 2 ceil(2^23 / n), add = 0. s * ceil(2^23 / n) / 2^23 exceeds s / n by less
//   than s / 2^23 <= 255 n / 2^23 < 1 / n (255 n^2 < 2^23), which cannot reach the next
//   integer, and s * mul <= 255 * 2^24 + 510 n < 2^32.
// - n = 1: mul = add = 2^24 - 1: s (2^24 - 1) + 2^24 - 1 = s 2^24 + (2^24 - 1 - s).
// - n = 0: mul = 0, add = 255 << 24 (the window sum is 0).
__host__ __device__ inline u32x2 sad_div_entry(uint32_t n) {
    if (n == 0) return u32x2{0u, 0xFF000000u};
    if (n == 1) return u32x2{0xFFFFFFu, 0xFFFFFFu};
    return u32x2{2u * (((1u << 23) + n - 1) / n), 0u};
}

// Dword offset of left-tile column c in a run segment: one column of nq dwords after another
// with a gap of nq dwords every 16 columns, so the (up to four) runs a wave reads at a step,
// 16 columns apart, fall on different banks.
__host__ __device__ inline int sad_lcol(int c, int nq) { return nq * (c + (c >> 4)); }
// Right-tile columns of a run segment sit in four planes by column index mod 4: the threads of
// a run read columns 4 apart at each step, i.e. consecutive columns of one plane.
__host__ __device__ inline int sad_plane_dwords(int tr, int nq) { return nq * ((tr + 3) / 4); }

// SAD segment LDS: prefiltered left/right tiles stored column-major (the 2r+1 rows of a column
// packed four per dword); the run segment adds a table of 8 divisors (eight-byte aligned), the
// generic segment its output tile [kSadTile][Dp].
__host__ __device__ inline size_t sad_lds_bytes(int r, int Dp) {
    const int nq = (2 * r + 1 + 3) / 4;
    const int P = sad_segment_px(r, Dp), tl = P + 2 * r, tr = tl + Dp - 1;
    if (sad_runs_radius(r))
        return (size_t)(sad_lcol(tl - 1, nq) + nq + 4 * sad_plane_dwords(tr, nq) + 17) * 4;
    return (size_t)tl * nq * 4 + (size_t)tr * nq * 4 + (size_t)P * Dp;
}
constexpr size_t kSadMaxLds = 64 * 1024;

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

// Block-wide; every thread of the block must call it with the same (y, x0). T is the segment
// width (a schedule choice; the costs do not depend on it).
//
// After the codes are in LDS the segment is produced in passes of 64 pixels: lane i of a wave
// takes pixel p0 + i and the wave's 16-disparity chunk c, so the right codes the 64 lanes read
// at each step are 64 consecutive words (no bank conflicts); the results go to the pass's LDS
// tile, which the block then stores as one contiguous run of 16-byte pieces.
template <int T = kCensusTile>
__device__ __forceinline__ void census_cost_segment(const PairJob& j, int y, int x0, uint32_t* shc) {
    const int W = j.W, H = j.H, D = j.D, Dp = j.Dp, m = j.m;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int rs = x0 - m - (Dp - 1);  // first right pixel the segment can reach
    const int nr = T + Dp - 1;         // right pixels the segment can reach
    const int lpitch = T + 4, rpitch = nr + 4, opitch = Dp + 16;
    uint8_t* tile = reinterpret_cast<uint8_t*>(shc);                       // 64 x opitch bytes
    uint32_t* sl = reinterpret_cast<uint32_t*>(tile + kCensusPass * opitch);  // T codes
    uint32_t* sr = sl + T;                                                  // nr codes
    uint8_t* il = reinterpret_cast<uint8_t*>(sr + nr);                      // 5 x lpitch bytes
    uint8_t* ir = il + 5 * lpitch;                                          // 5 x rpitch bytes
    const bool rowOk = y >= 2 && y < H - 2;
    if (rowOk) {
        // One staged column (five rows) per thread and iteration, all five loads issued before
        // the first use; columns outside the image stage zeros (their loads read a clamped,
        // in-image column).
        const LUMEN_G uint8_t* l0 = gmem(j.left) + (size_t)(y - 2) * W;
        const LUMEN_G uint8_t* r0 = gmem(j.right) + (size_t)(y - 2) * W;
        for (int i = tid; i < lpitch + rpitch; i += nt) {
            const bool isL = i < lpitch;
            const int c = isL ? x0 - 2 + i : rs - 2 + (i - lpitch);
            const bool ok = c >= 0 && c < W;
            const LUMEN_G uint8_t* src = (isL ? l0 : r0) + min(max(c, 0), W - 1);
            uint8_t* dst = isL ? il + i : ir + (i - lpitch);
            const int pitch = isL ? lpitch : rpitch;
            uint8_t v[5];
#pragma unroll
            for (int r = 0; r < 5; ++r) v[r] = src[(size_t)r * W];
#pragma unroll
            for (int r = 0; r < 5; ++r) dst[r * pitch] = ok ? v[r] : uint8_t(0);
        }
        __syncthreads();
        for (int i = tid; i < T; i += nt) {
            const int x = x0 + i;
            sl[i] = (x >= 2 && x < W - 2) ? census_at(il, lpitch, i + 2) : 0u;
        }
        for (int i = tid; i < nr; i += nt) {
            const int rx = rs + i;
            sr[i] = (rx >= 2 && rx < W - 2) ? census_at(ir, rpitch, i + 2) : 0u;
        }
        __syncthreads();
    }
    const int chunks = Dp / 16;
    const int lane = tid & 63, wave = tid >> 6, nwaves = nt >> 6;
    // Copy-out: thread tid keeps 16-byte chunk cc of every per-th pixel of the pass.
    const int per = nt / chunks, cc = tid % chunks, pp = tid / chunks;
    for (int p0 = 0; p0 < T && x0 + p0 < W; p0 += kCensusPass) {
        const int i = p0 + lane;
        const int x = x0 + i;
        const bool leftOk = rowOk && x >= 2 && x < W - 2;
        const uint32_t a = leftOk ? sl[i] : 0u;
        for (int c = wave; c < chunks; c += nwaves) {
            const int d0 = c * 16;
            // rc[-k] is the right code for d0 + k; index i + Dp - 1 - d0 - k stays in [0, nr)
            // for every lane and k, so it is read even where the result is then discarded.
            const uint32_t* rc = sr + i + Dp - 1 - d0;
            uint32_t w[4];
            // (h*255+12)/24 == (h*2718+168) >> 8 for h in [0,24]; two results per packed
            // 16-bit multiply-add, byte 1 of each half gathered with one byte permute per four.
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
            const int e = x - m - d0;  // right column of disparity d0
            if (!(leftOk && d0 + 15 < D && e - 15 >= 2 && e < W - 2)) {
                // Not every disparity of the chunk is valid. The valid k (d = d0 + k) form one
                // run: d < D, 2 <= x - m - d < W - 2; the others read 255.
                const int klo = max(0, e - (W - 3)), khi = min(min(15, D - 1 - d0), e - 2);
                const uint32_t valid =
                    (leftOk && khi >= klo) ? (((2u << khi) - 1u) & ~((1u << klo) - 1u)) : 0u;
                const uint32_t inv = ~valid;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    // Nibble bit b -> byte b (bit b lands at 8b through the 7b shift), then 0xFF.
                    const uint32_t n = (inv >> (4 * q)) & 0xFu;
                    w[q] |= ((n * 0x00204081u) & 0x01010101u) * 0xFFu;
                }
            }
            *reinterpret_cast<u32x4*>(tile + lane * opitch + d0) = u32x4{w[0], w[1], w[2], w[3]};
        }
        __syncthreads();
        const int npx = min(kCensusPass, W - (x0 + p0));
        LUMEN_G u32x4* dst =
            reinterpret_cast<LUMEN_G u32x4*>(gmem(j.cost) + ((size_t)y * W + x0 + p0) * Dp);
        if (pp < per)
            for (int pi = pp; pi < npx; pi += per)
                dst[pi * chunks + cc] = *reinterpret_cast<const u32x4*>(tile + pi * opitch + cc * 16);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// Sobel-x prefilter (src/sgbm_kernels.cu:38-59) fused into the SAD cost
// (src/sgbm_kernels.cu:172-212): one row segment of kSadTile pixels per call, block-wide. The
// prefiltered rows y-r..y+r of the segment and of the right columns it can reach are built in
// LDS. The valid taps of a window form a rectangle, so count = rows * columns.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ uint8_t prefilter_at(const LUMEN_G uint8_t* __restrict__ row, int c, int W,
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

// Four prefiltered bytes (src/sgbm_kernels.cu:38-59) from the bytes b0..b7 of one image row
// held in lo (b0..b3) and hi (b4..b7), b0 being the column left of the first: byte j is
// clamp(b[j+2] - b[j], -cap, cap) + cap mod 256, computed as two packed 16-bit lanes (even and
// odd j). clampv = min(cap, 255) clamps the same way (|difference| <= 255), addv = cap mod 256
// gives the same low byte.
__device__ __forceinline__ uint32_t prefilter4(uint32_t lo, uint32_t hi, i16x2 clampv, i16x2 addv) {
    const uint32_t mn = __builtin_amdgcn_alignbyte(hi, lo, 2);  // b2..b5
    const i16x2 m02 = __builtin_bit_cast(i16x2, __builtin_amdgcn_perm(0u, mn, 0x0c020c00u));
    const i16x2 m13 = __builtin_bit_cast(i16x2, __builtin_amdgcn_perm(0u, mn, 0x0c030c01u));
    const i16x2 s02 = __builtin_bit_cast(i16x2, __builtin_amdgcn_perm(0u, lo, 0x0c020c00u));
    const i16x2 s13 = __builtin_bit_cast(i16x2, __builtin_amdgcn_perm(0u, lo, 0x0c030c01u));
    const i16x2 v02 = __builtin_elementwise_min(__builtin_elementwise_max(m02 - s02, -clampv), clampv) + addv;
    const i16x2 v13 = __builtin_elementwise_min(__builtin_elementwise_max(m13 - s13, -clampv), clampv) + addv;
    // Low bytes of the four halves, in column order.
    return __builtin_amdgcn_perm(__builtin_bit_cast(uint32_t, v13), __builtin_bit_cast(uint32_t, v02),
                                 0x06020400u);
}

// Block-wide; every thread of the block must call it with the same (y, x0): stages the
// prefiltered tiles of a radius-R segment (R = 1..3). Left-tile column c (c < TL) is image
// column lbase + c, right-tile column k (k < TR) image column rbase + k; a column's 2R+1 rows
// are packed four per dword, at lt + sad_lcol(c, NQ) and at rt + (k & 3) * PD + NQ * (k >> 2).
// Rows outside the image stage 0.
//
// sad_stage_columns: any segment. CPT columns per thread and pass, all their rows' neighbour
// pairs loaded before the first use (clamped indices, so every load is inside the image): a
// segment of up to CPT * 256 columns costs one round trip to memory. Clamping the neighbours
// reproduces prefilter_at's edge rule: column 0 takes row[1] - row[0], column W-1 row[W-1] -
// row[W-2]. Taps outside the image stage 0 in both tiles.
template <int R, int CPT>
__device__ __forceinline__ void sad_stage_columns(const PairJob& j, int y, int TL, int TR,
                                                  int lbase, int rbase, int cap, uint32_t* lt,
                                                  uint32_t* rt, int PD) {
    constexpr int NQ = (2 * R + 1 + 3) / 4;
    const int W = j.W, H = j.H;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int cb = tid; cb < TL + TR; cb += CPT * nt) {
        uint32_t a[CPT][2 * R + 1], b[CPT][2 * R + 1];
#pragma unroll
        for (int h = 0; h < CPT; ++h) {
            const int c = min(cb + h * nt, TL + TR - 1);  // a past-the-end column loads a real one
            const bool isL = c < TL;
            const int gc = isL ? lbase + c : rbase + (c - TL);
            const int cl = min(max(gc - 1, 0), W - 1), cr = min(max(gc + 1, 0), W - 1);
            const LUMEN_G uint8_t* img = gmem(isL ? j.left : j.right);
#pragma unroll
            for (int k = 0; k <= 2 * R; ++k) {
                const size_t row = (size_t)min(max(y - R + k, 0), H - 1) * W;
                a[h][k] = img[row + cr];
                b[h][k] = img[row + cl];
            }
        }
#pragma unroll
        for (int h = 0; h < CPT; ++h) {
            const int c = cb + h * nt;
            if (c >= TL + TR) break;
            const int gc = c < TL ? lbase + c : rbase + (c - TL);
            const bool colOk = gc >= 0 && gc < W;
            uint32_t w[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) w[q] = 0;
#pragma unroll
            for (int k = 0; k <= 2 * R; ++k) {
                const int yy = y - R + k;
                const int g = min(cap, max(-cap, static_cast<int>(a[h][k]) - static_cast<int>(b[h][k])));
                // The prefiltered value is stored as uint8 (src/sgbm_kernels.cu:58): for caps of
                // 128 and more g + cap can exceed 255 and wraps mod 256, and must not spill into
                // the next row's byte.
                const uint32_t v =
                    (colOk && yy >= 0 && yy < H) ? (static_cast<uint32_t>(g + cap) & 0xFFu) : 0u;
                w[k / 4] |= v << (8 * (k % 4));
            }
            const int k = c - TL;
            uint32_t* out = c < TL ? lt + sad_lcol(c, NQ) : rt + (k & 3) * PD + NQ * (k >> 2);
#pragma unroll
            for (int q = 0; q < NQ; ++q) out[q] = w[q];
        }
    }
}

// sad_stage_groups: segments whose every source column lies inside the image, twelve bytes
// before its row's end (sad_groups_inside), in 4-byte aligned images. Each thread takes a
// group of four consecutive columns of one tile and loads each of its rows as the three aligned
// dwords around image columns gc0-1 .. gc0+4 (one round trip for segments of up to 1024
// columns), forms the four prefiltered bytes with byte selects and packed 16-bit arithmetic
// (prefilter4), transposes rows into columns with byte permutes and stores the columns inside
// its tile (the last group of a tile may be partial): about a third of the per-column form's
// instructions and an eighth of its loads.
__host__ __device__ inline bool sad_groups_inside(int TL, int TR, int lbase, int rbase, int W) {
    return lbase >= 1 && lbase + 4 * ((TL + 3) / 4 - 1) + 11 <= W && rbase >= 1 &&
           rbase + 4 * ((TR + 3) / 4 - 1) + 11 <= W;
}

template <int R>
__device__ __forceinline__ void sad_stage_groups(const PairJob& j, int y, int TL, int TR,
                                                 int lbase, int rbase, int cap, uint32_t* lt,
                                                 uint32_t* rt, int PD) {
    constexpr int NQ = (2 * R + 1 + 3) / 4, NR = 2 * R + 1;
    const int W = j.W, H = j.H;
    const int GL = (TL + 3) / 4, GR = (TR + 3) / 4;
    const i16x2 clampv = static_cast<short>(min(cap, 255));
    const i16x2 addv = static_cast<short>(cap & 0xFF);
    uint32_t rowOff[NR], rowMask[NR];  // rows outside the image: a clamped row, masked to 0
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const int yy = y - R + k;
        rowMask[k] = yy >= 0 && yy < H ? ~0u : 0u;
        rowOff[k] = static_cast<uint32_t>(min(max(yy, 0), H - 1)) * static_cast<uint32_t>(W);
    }
    for (int gi = threadIdx.x; gi < GL + GR; gi += blockDim.x) {
        const bool isL = gi < GL;
        const int c0 = 4 * (isL ? gi : gi - GL);  // first column of the group in its tile
        const int n = isL ? TL : TR;
        const int gc0 = (isL ? lbase : rbase) + c0;
        const LUMEN_G uint8_t* img = gmem(isL ? j.left : j.right);
        // Column c0 + t of the group is at col0 + t * cstep (the four columns of a left group
        // share one 16-column run, those of a right group one index in the four planes).
        uint32_t* const col0 = isL ? lt + sad_lcol(c0, NQ) : rt + NQ * (c0 >> 2);
        const int cstep = isL ? NQ : PD;
        u32x3 w[NR];
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const uint32_t a = rowOff[k] + static_cast<uint32_t>(gc0 - 1);
            w[k] = *reinterpret_cast<const LUMEN_G u32x3*>(img + (a & ~3u));
        }
        uint32_t P[4 * NQ];
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const uint32_t off = (rowOff[k] + static_cast<uint32_t>(gc0 - 1)) & 3u;
            const uint32_t lo = __builtin_amdgcn_alignbyte(w[k].y, w[k].x, off);
            const uint32_t hi = __builtin_amdgcn_alignbyte(w[k].z, w[k].y, off);
            P[k] = prefilter4(lo, hi, clampv, addv) & rowMask[k];
        }
#pragma unroll
        for (int k = NR; k < 4 * NQ; ++k) P[k] = 0u;
        // 4x4 byte transposes: dword q of column t holds rows 4q .. 4q+3 of that column.
        uint32_t col[4][NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const uint32_t A = P[4 * q], B = P[4 * q + 1], C = P[4 * q + 2], Dd = P[4 * q + 3];
            const uint32_t abLo = __builtin_amdgcn_perm(B, A, 0x05010400u);
            const uint32_t abHi = __builtin_amdgcn_perm(B, A, 0x07030602u);
            const uint32_t cdLo = __builtin_amdgcn_perm(Dd, C, 0x05010400u);
            const uint32_t cdHi = __builtin_amdgcn_perm(Dd, C, 0x07030602u);
            col[0][q] = __builtin_amdgcn_perm(cdLo, abLo, 0x05040100u);
            col[1][q] = __builtin_amdgcn_perm(cdLo, abLo, 0x07060302u);
            col[2][q] = __builtin_amdgcn_perm(cdHi, abHi, 0x05040100u);
            col[3][q] = __builtin_amdgcn_perm(cdHi, abHi, 0x07060302u);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            if (c0 + t < n) {
#pragma unroll
                for (int q = 0; q < NQ; ++q) col0[t * cstep + q] = col[t][q];
            }
        }
    }
}

// Window costs of radius R = 1..3 (block sizes 3, 5, 7; src/sgbm_kernels.cu:172-212 with the
// Sobel-x prefilter of :38-59 fused in) for one row segment of P = sad_segment_px(R, Dp)
// pixels, block-wide; every thread of the block must call it with the same (y, x0).
//
// Thread (q, part) owns disparities 4q .. 4q+3 and the kSadRun pixels of its part. For each
// disparity it keeps the running sum P(c) of the column SADs col(c) (v_sad_u8 over the 2R+1
// staged rows of left column c and the right column c - (m+d)) in a ring of the last 2R+1
// values: the window of pixel i, left columns i .. i+2R, is P(i+2R) - P(i-1). A column whose
// right column lies left of the image, or which lies right of it, adds nothing; the valid taps
// of a window form a rectangle, so their count is rows * columns. A run whose windows all span
// the full 2R+1 columns skips those checks and divides by the block's window size n (2R+1
// columns by the rows inside the image) with a multiply: floor(s / n) is byte 3 of
// s * ceil(2^24 / n) for s <= 255 n and 3 <= n <= 49 (255 n^2 <= 2^24, ceil(2^24 / n) < 2^24,
// and 255 n * ceil(2^24 / n) < 2^32). The four costs of a pixel are one dword of the volume.
template <int R>
__device__ __forceinline__ void sad_cost_runs(const PairJob& j, int y, int x0, uint32_t* lds) {
    constexpr int NQ = (2 * R + 1 + 3) / 4, NW = 2 * R + 1;
    const int W = j.W, H = j.H, D = j.D, Dp = j.Dp, m = j.m;
    const int cap = j.preFilterCap > 1 ? j.preFilterCap : 1;
    const int runs = 1024 / Dp, P = kSadRun * runs;
    const int TL = P + 2 * R, TR = TL + Dp - 1;
    const int PD = sad_plane_dwords(TR, NQ);
    uint32_t* const lt = lds;
    uint32_t* const rt = lds + sad_lcol(TL - 1, NQ) + NQ;
    const int dylo = max(-R, -y), dyhi = min(R, H - 1 - y);
    const int nrows = dyhi - dylo + 1;  // rows of this row's windows inside the image
    // Divisors of the checked windows: div[c] = (mul, add) for c valid columns, n = c * nrows
    // taps, with floor(s / n) = byte 3 of s * mul + add for every window sum s <= 255 n
    // (sad_div_entry); a window without valid columns (s = 0) reads 255.
    u32x2* const div = reinterpret_cast<u32x2*>(rt + 4 * PD + ((sad_lcol(TL - 1, NQ) + NQ) & 1));
    if (threadIdx.x < 8) div[threadIdx.x] = sad_div_entry(threadIdx.x * nrows);
    // Segments away from the image borders stage by column groups (at most 148 groups of four,
    // for Dp = 64 or 512 with R = 3: one pass), the others column by column (up to 768 columns,
    // Dp up to 448 with R = 3, in one pass). The choice is the block's, so no wave runs both.
    const int lbase = x0 - R, rbase = x0 - R - m - (Dp - 1);
    const bool aligned =
        ((reinterpret_cast<uintptr_t>(j.left) | reinterpret_cast<uintptr_t>(j.right)) & 3u) == 0;
    if (aligned && sad_groups_inside(TL, TR, lbase, rbase, W))
        sad_stage_groups<R>(j, y, TL, TR, lbase, rbase, cap, lt, rt, PD);
    else
        sad_stage_columns<R, 3>(j, y, TL, TR, lbase, rbase, cap, lt, rt, PD);
    __syncthreads();
    const int quads = Dp / 4, tid = threadIdx.x;
    const int q = tid % quads, part = tid / quads;
    const int i0 = part * kSadRun;
    const int npx = min(P, W - x0);
    if (part >= runs || i0 >= npx) return;
    const int n = min(kSadRun, npx - i0);
    const int dq = 4 * q;
    LUMEN_G uint32_t* const out = reinterpret_cast<LUMEN_G uint32_t*>(
        gmem(j.cost) + ((size_t)y * W + x0 + i0) * Dp + dq);
    const int ostep = Dp / 4;  // dwords from one pixel's costs to the next's
    if (dq >= D) {
        for (int i = 0; i < n; ++i) out[i * ostep] = ~0u;
        return;
    }
    const uint32_t padMask = D - dq >= 4 ? 0u : (~0u << (8 * (D - dq)));  // d >= D reads 255
    // Left column i0 + t is at lb + NQ (t + t/16) (i0 is a multiple of 16). Right column
    // i0 + t + Dp - 1 - dq - k (the one disparity dq + k pairs with left column i0 + t) is
    // 4 bq + 3 + t - k, which is plane (3 + t - k) & 3 at rb + NQ ((3 + t - k) >> 2).
    const uint32_t* const lb = lt + NQ * (i0 + part);
    const int bq = i0 / 4 + quads - 1 - q;
    auto lcol = [&](int t, uint32_t(&v)[NQ]) {
        const uint32_t* p = lb + NQ * (t + (t >> 4));
        if constexpr (NQ == 2) {
            const u32x2 w = *reinterpret_cast<const u32x2*>(p);
            v[0] = w.x;
            v[1] = w.y;
        } else {
            v[0] = p[0];
        }
    };
    auto rcol = [&](int e, uint32_t(&v)[NQ]) {  // e = 3 + t - k >= 0
        const uint32_t* p = rt + (e & 3) * PD + NQ * (bq + (e >> 2));
        if constexpr (NQ == 2) {
            const u32x2 w = *reinterpret_cast<const u32x2*>(p);
            v[0] = w.x;
            v[1] = w.y;
        } else {
            v[0] = p[0];
        }
    };
    // rr[k]: the right column disparity dq + k pairs with the current left column. Before
    // column t = 0, rr[0..2] hold those of column -1 (shifted into rr[1..3] at t = 0).
    uint32_t rr[4][NQ];
#pragma unroll
    for (int k = 0; k < 3; ++k) rcol(2 - k, rr[k]);
    const int xf = x0 + i0;
    const int tLo = m + dq - (xf - R), tHi = W - (xf - R);
    const uint32_t nwin = static_cast<uint32_t>(NW * nrows);
    const uint32_t magic = ((1u << 24) + nwin - 1) / nwin;
    // The checked form is right for every run; a wave takes it as a whole when any of its runs
    // needs it (a wave split between the two forms would execute both). Such waves take
    // longer: they issue first within their SIMD.
    const bool whole =
        __all(n == kSadRun && xf - R >= m + dq + 3 && xf + kSadRun - 1 + R <= W - 1);
    if (!whole) __builtin_amdgcn_s_setprio(2);
    uint32_t acc[4] = {0u, 0u, 0u, 0u};
    auto enter = [&](int t, bool checked) {
        // Left column t enters: rr shifts by one column, then each disparity's running sum
        // takes the column's SAD (in a checked run, only when the column is a valid tap).
#pragma unroll
        for (int k = 3; k > 0; --k)
#pragma unroll
            for (int e = 0; e < NQ; ++e) rr[k][e] = rr[k - 1][e];
        rcol(3 + t, rr[0]);
        uint32_t lv[NQ];
        lcol(t, lv);
        // Left column t (image column xf - R + t) pairs with right image column
        // xf - R + t - m - dq - k for disparity dq + k: a valid tap when t - k >= tLo and
        // t < tHi (t - k is a constant in the unrolled runs).
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (checked) {
                uint32_t a = acc[k];
#pragma unroll
                for (int e = 0; e < NQ; ++e) a = __builtin_amdgcn_sad_u8(lv[e], rr[k][e], a);
                // Both tests evaluated (no short circuit), so the select stays a select.
                acc[k] = ((t - k >= tLo) & (t < tHi)) ? a : acc[k];
            } else {
#pragma unroll
                for (int e = 0; e < NQ; ++e) acc[k] = __builtin_amdgcn_sad_u8(lv[e], rr[k][e], acc[k]);
            }
        }
    };
    auto emit = [&](int i, const uint32_t(&p)[4]) {  // p: each cost in byte 3
        const uint32_t lo = __builtin_amdgcn_perm(p[1], p[0], 0x0c0c0703u);
        const uint32_t hi = __builtin_amdgcn_perm(p[3], p[2], 0x0c0c0703u);
        out[i * ostep] = __builtin_amdgcn_perm(hi, lo, 0x05040100u) | padMask;
    };
    // Costs (in byte 3) of pixel x's windows whose taps were checked column by column: the valid
    // columns and rows form a rectangle, so the divisor follows from the valid column count. No
    // branches: empty windows read their table entry too.
    auto checked_costs = [&](int x, const uint32_t(&sum)[4], uint32_t(&p)[4]) {
        const int hi = min(R, W - 1 - x) + 1, b = m + dq - x;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const u32x2 dv = div[max(hi - max(-R, b + k), 0)];
            p[k] = __umul24(sum[k], dv.x) + dv.y;
        }
    };
    if (n == kSadRun) {
        // Fully unrolled: the ring slots are registers, the LDS offsets immediates, and the
        // steps have no branches, so the scheduler overlaps one step's loads with the last
        // one's arithmetic. A run whose windows are all whole skips the checks.
        auto run = [&](auto checked) {
            constexpr bool CHECKED = decltype(checked)::value;
            uint32_t ring[4][NW];
#pragma unroll
            for (int k = 0; k < 4; ++k) ring[k][NW - 1] = 0u;  // P(i0 - 1); the others fill first
            static_for<0, kSadRun + 2 * R>([&](auto tc) {
                constexpr int t = decltype(tc)::value;
                enter(t, CHECKED);
                if constexpr (t >= 2 * R) {
                    uint32_t p[4];
                    if constexpr (CHECKED) {
                        uint32_t sum[4];
#pragma unroll
                        for (int k = 0; k < 4; ++k) sum[k] = acc[k] - ring[k][t % NW];
                        checked_costs(xf + t - 2 * R, sum, p);
                    } else {
#pragma unroll
                        for (int k = 0; k < 4; ++k) p[k] = __umul24(acc[k] - ring[k][t % NW], magic);
                    }
                    emit(t - 2 * R, p);
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) ring[k][t % NW] = acc[k];
            });
        };
        if (whole)
            run(std::integral_constant<bool, false>{});
        else
            run(std::integral_constant<bool, true>{});
    } else {
        // The short run at the end of a row whose width is not a multiple of 16, rolled.
        // hist[k] holds P(c - 2R - 1) .. P(c - 1), rotated every step. A step whose lanes all
        // enter a valid column and emit a window of whole rows' width skips the per-column
        // checks and divides by the block's window size (fewer rows at the top and bottom of
        // the image).
        uint32_t hist[4][NW];
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int e = 0; e < NW; ++e) hist[k][e] = 0u;
#pragma unroll 1
        for (int t = 0; t < n + 2 * R; ++t) {
            const int lo = xf - R + t - (t >= 2 * R ? 2 * R : 0);  // leftmost column needed
            const bool fast = __all(xf - R + t <= W - 1 && lo >= m + dq + 3);
            enter(t, !fast);
            if (t >= 2 * R) {
                uint32_t sum[4], p[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) sum[k] = acc[k] - hist[k][0];
                if (fast) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) p[k] = __umul24(sum[k], magic);
                } else {
                    checked_costs(xf + t - 2 * R, sum, p);
                }
                emit(t - 2 * R, p);
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
#pragma unroll
                for (int e = 0; e < NW - 1; ++e) hist[k][e] = hist[k][e + 1];
                hist[k][NW - 1] = acc[k];
            }
        }
    }
}

// Block-wide; every thread of the block must call it with the same (y, x0); the segment's LDS
// (sad_lds_bytes) must fit kSadMaxLds. Any radius (the run segment covers 1 to 3): one row
// segment of kSadTile pixels, the prefiltered tiles built byte by byte.
//
// For disparity d the window sum is a sliding sum over columns: col(c) is the SAD of the 2r+1
// staged rows of left column c against right column c - (m+d), and zero when either column is
// outside the image (its taps do not count). A thread owns one disparity and one part of the
// segment and slides along x, recomputing the column that leaves.
__device__ __forceinline__ void sad_cost_segment_generic(const PairJob& j, int y, int x0, uint8_t* shs) {
    const int W = j.W, H = j.H, D = j.D, Dp = j.Dp, m = j.m;
    const int r = j.blockSize / 2;
    const int cap = j.preFilterCap > 1 ? j.preFilterCap : 1;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int rows = 2 * r + 1, nq = (rows + 3) / 4;
    constexpr int T = kSadTile;
    const int TL = T + 2 * r, TR = T + 2 * r + Dp - 1;
    uint32_t* lt = reinterpret_cast<uint32_t*>(shs);  // TL columns x nq dwords
    uint32_t* rt = lt + TL * nq;                          // TR columns x nq dwords
    uint8_t* ot = reinterpret_cast<uint8_t*>(rt + TR * nq);  // [T][Dp]
    const int lbase = x0 - r, rbase = x0 - r - m - (Dp - 1);
    uint8_t* lb = reinterpret_cast<uint8_t*>(lt);
    uint8_t* rb = reinterpret_cast<uint8_t*>(rt);
    for (int i = tid; i < TL * nq * 4; i += nt) {
        const int c = i / (nq * 4), k = i - c * (nq * 4);
        const int gc = lbase + c, yy = y - r + k;
        lb[i] = (k < rows && yy >= 0 && yy < H && gc >= 0 && gc < W)
                    ? prefilter_at(gmem(j.left) + (size_t)yy * W, gc, W, cap) : uint8_t(0);
    }
    for (int i = tid; i < TR * nq * 4; i += nt) {
        const int c = i / (nq * 4), k = i - c * (nq * 4);
        const int gc = rbase + c, yy = y - r + k;
        rb[i] = (k < rows && yy >= 0 && yy < H && gc >= 0 && gc < W)
                    ? prefilter_at(gmem(j.right) + (size_t)yy * W, gc, W, cap) : uint8_t(0);
    }
    __syncthreads();
    const int dylo = max(-r, -y), dyhi = min(r, H - 1 - y);
    const int nrows = dyhi - dylo + 1;
    const uint32_t full = static_cast<uint32_t>(rows * rows);
    // floor(s / full) == umulhi(s, ceil(2^32 / full)) for s <= 255 * full (255 * full^2 < 2^32).
    // A 1x1 window (full = 1) has no 32-bit multiplier: its sum is the cost itself.
    const uint32_t magic = static_cast<uint32_t>((0x100000000ull + full - 1) / full);
    const int parts = max(1, nt / Dp);
    const int plen = (T + parts - 1) / parts;
    for (int t = tid; t < Dp * parts; t += nt) {
        const int d = t % Dp, part = t / Dp;
        const int i0 = part * plen, i1 = min(T, i0 + plen);
        if (i0 >= i1) continue;
        if (d >= D) {
            for (int i = i0; i < i1; ++i) ot[i * Dp + d] = 255;
            continue;
        }
        const int a = m + d;
        // col(c): left tile column c (image column lbase + c) against right tile column
        // c + Dp - 1 - d (image column lbase + c - a).
        auto col = [&](int c) -> uint32_t {
            const int gc = lbase + c;
            if (gc < a || gc >= W) return 0u;
            const uint32_t* L = lt + c * nq;
            const uint32_t* Rr = rt + (c + Dp - 1 - d) * nq;
            uint32_t acc = 0;
            for (int q = 0; q < nq; ++q) acc = __builtin_amdgcn_sad_u8(L[q], Rr[q], acc);
            return acc;
        };
        auto emit = [&](int i, uint32_t sum) {
            const int x = x0 + i;
            const int dxlo = max(-r, a - x), dxhi = min(r, W - 1 - x);
            const int ncols = dxhi - dxlo + 1;
            uint32_t v;
            if (ncols == rows && nrows == rows)
                v = full == 1 ? sum : __umulhi(sum, magic);
            else if (ncols > 0 && nrows > 0)
                v = sum / static_cast<uint32_t>(ncols * nrows);
            else
                v = 255;
            ot[i * Dp + d] = static_cast<uint8_t>(v);
        };
        // Window of output pixel i covers tile columns i .. i + 2r.
        uint32_t sum = 0;
        for (int k = 0; k < 2 * r; ++k) sum += col(i0 + k);
        for (int i = i0; i < i1; ++i) {
            sum += col(i + 2 * r);
            emit(i, sum);
            sum -= col(i);
        }
    }
    __syncthreads();
    // The segment's costs are contiguous in the volume: 16-byte stores.
    const int npix = min(T, W - x0);
    const int n16 = npix * Dp / 16;
    LUMEN_G u32x4* dst = reinterpret_cast<LUMEN_G u32x4*>(gmem(j.cost) + ((size_t)y * W + x0) * Dp);
    const u32x4* src = reinterpret_cast<const u32x4*>(ot);
    for (int i = tid; i < n16; i += nt) dst[i] = src[i];
}

// One row segment of sad_segment_px(blockSize / 2, Dp) pixels starting at x0.
__device__ __forceinline__ void sad_cost_segment(const PairJob& j, int y, int x0, uint8_t* shs) {
    switch (j.blockSize / 2) {
        case 1: sad_cost_runs<1>(j, y, x0, reinterpret_cast<uint32_t*>(shs)); break;
        case 2: sad_cost_runs<2>(j, y, x0, reinterpret_cast<uint32_t*>(shs)); break;
        case 3: sad_cost_runs<3>(j, y, x0, reinterpret_cast<uint32_t*>(shs)); break;
        default: sad_cost_segment_generic(j, y, x0, shs); break;
    }
}

// SAD cost of one (pixel, disparity) read from the images directly (windows whose segment does
// not fit in LDS).
__device__ __forceinline__ uint8_t sad_cost_direct(const PairJob& j, int x, int y, int d) {
    const int W = j.W, H = j.H;
    const int r = j.blockSize / 2;
    const int cap = j.preFilterCap > 1 ? j.preFilterCap : 1;
    if (d >= j.D) return 255;
    const int a = j.m + d;
    const int dylo = max(-r, -y), dyhi = min(r, H - 1 - y);
    const int dxlo = max(-r, a - x), dxhi = min(r, W - 1 - x);
    if (dyhi < dylo || dxhi < dxlo) return 255;
    int sum = 0;
    for (int dy = dylo; dy <= dyhi; ++dy) {
        const LUMEN_G uint8_t* lr = gmem(j.left) + (size_t)(y + dy) * W;
        const LUMEN_G uint8_t* rr = gmem(j.right) + (size_t)(y + dy) * W;
        for (int dx = dxlo; dx <= dxhi; ++dx)
            sum += abs(int(prefilter_at(lr, x + dx, W, cap)) - int(prefilter_at(rr, x + dx - a, W, cap)));
    }
    return static_cast<uint8_t>(sum / ((dyhi - dylo + 1) * (dxhi - dxlo + 1)));
}

// ---------------------------------------------------------------------------------------
// SGM aggregation, one wave = four neighbouring scanlines of one path.
//
// A scanline is the sequence of pixels of one path from its origin (the pixel whose
// predecessor lies outside the image) to the far border. Each scanline is owned by one DPP row
// (16 lanes); lane l holds disparities [l*K, l*K+K) as K/2 packed uint16 pairs, so the DP step
// is packed 16-bit VALU work, the d-1 / d+1 neighbours cross lanes with one row_shr/row_shl
// each, and min over d is a 4-step DPP butterfly. Scanline geometry is origin + t*step in
// flat pixel index. Path order: 0 LR, 1 RL, 2 TB, 3 BT, 4 TLBR, 5 BRTL, 6 TRBL, 7 BLTR.
// ---------------------------------------------------------------------------------------
__host__ __device__ inline int scan_count(int path, int W, int H) {
    return path < 2 ? H : (path < 4 ? W : W + H - 1);
}

// Length of scanline s of a path (0 when s is past the last scanline).
__host__ __device__ inline int scan_length(int path, int s, int W, int H) {
    if (s >= scan_count(path, W, H)) return 0;
    if (path < 2) return W;
    if (path < 4) return H;
    return s < H - 1 ? min(W, s + 1) : min(W - (s - (H - 1)), H);
}

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

// Slice layout. A scanline's disparity vector is spread over LPS lanes (a 16-lane DPP row, 32
// lanes: two rows, or the whole 64-lane wave), lane l holding K of the pixel's Dp = LPS*K
// disparities.
//
// LPS = 16: K/Q chunks of Q consecutive disparities (Q = 8 when K is a multiple of 8, 4 when
// it is a multiple of 4, else K: K = 6, Dp = 96): chunk c of lane l is d = 16Qc + Ql ..
// 16Qc + Ql + Q-1. A chunk is one 4- or 8-byte cost load (three 2-byte ones for K = 6) and one
// 8-, 12- or 16-byte path-cost access per lane, and the 16 lanes' chunks are contiguous, so
// every access instruction of a row covers contiguous bytes whatever K is. For K = 4, 6 and 8
// there is one chunk and lane l simply holds d = lK .. lK+K-1.
// LPS = 32 or 64: for K = 2, 4, 8 or 16 one chunk, lane l holding d = lK .. lK+K-1; otherwise
// chunks of Q = 4 (K = 12) or Q = 2 (K = 6, 10, 14) disparities, the LPS lanes of a chunk again
// covering LPS*Q consecutive disparities, so each access of a scanline covers 64 to 1024
// contiguous bytes. The chunks alternate in direction (snake_chunks): even chunks run up the
// lanes (lane l holds LPS*Qc + Ql ..), odd chunks down them (lane l holds LPS*Qc + Q(LPS-1-l)
// ..), so the disparities either side of a chunk boundary sit in the same lane (the last lane
// after an even chunk, lane 0 after an odd one) and no value crosses between DPP rows there.
// Registers hold the lane's elements chunk by chunk, element e being slice_d<K, LPS>(l, e).
template <int K, int LPS = 16>
__host__ __device__ constexpr int slice_q() {
    return LPS >= 32 ? ((K == 2 || K == 4 || K == 8 || K == 16) ? K : (K % 4 == 0 ? 4 : 2))
                     : (K % 8 == 0 ? 8 : K % 4 == 0 ? 4 : K);
}

template <int K, int LPS = 16>
__host__ __device__ constexpr bool snake_chunks() {
    return LPS >= 32 && K / slice_q<K, LPS>() > 1;
}

template <int K, int LPS = 16>
__device__ __forceinline__ int slice_d(int l, int e) {
    constexpr int Q = slice_q<K, LPS>();
    const int c = e / Q;
    const int ll = snake_chunks<K, LPS>() && (c & 1) ? LPS - 1 - l : l;
    return LPS * Q * c + Q * ll + e % Q;
}

// Elements from lane l's position (Q*l) to its position in an odd chunk of a snake layout.
template <int K, int LPS = 16>
__device__ __forceinline__ int odd_chunk_offset(int l) {
    constexpr int Q = slice_q<K, LPS>();
    return snake_chunks<K, LPS>() ? Q * (LPS - 1 - 2 * l) : 0;
}

// Lane that holds disparity d, and its element index there (16-lane rows).
template <int K>
__device__ __forceinline__ int slice_lane(int d) {
    constexpr int Q = slice_q<K>();
    return (d % (16 * Q)) / Q;
}
template <int K>
__device__ __forceinline__ int slice_elem(int d) {
    constexpr int Q = slice_q<K>();
    return d / (16 * Q) * Q + d % Q;
}

// Whether a lane's cost bytes are held two to a word (low half): chunks of two, or the
// six-disparity lane of K = 6; otherwise four to a word.
template <int K, int LPS = 16>
__host__ __device__ constexpr bool cost_pairs() {
    return slice_q<K, LPS>() == 2 || slice_q<K, LPS>() == 6;
}

// A lane's K cost bytes.
template <int K, int LPS = 16>
struct CostBytes {
    uint32_t w[cost_pairs<K, LPS>() ? K / 2 : K / 4];
};

// p: the pixel's cost vector + Q*l (the lane's first chunk); odd: odd_chunk_offset(l).
template <int K, int LPS = 16>
__device__ __forceinline__ CostBytes<K, LPS> load_cost(const LUMEN_G uint8_t* p, int odd = 0) {
    constexpr int Q = slice_q<K, LPS>();
    CostBytes<K, LPS> c;
#pragma unroll
    for (int ch = 0; ch < K / Q; ++ch) {
        const LUMEN_G uint8_t* q = p + LPS * Q * ch + (snake_chunks<K, LPS>() && (ch & 1) ? odd : 0);
        if constexpr (Q == 2) {
            c.w[ch] = *reinterpret_cast<const LUMEN_G uint16_t*>(q);
        } else if constexpr (Q == 6) {
            // Six bytes at a 2-byte aligned address: the eight bytes from the 4-byte aligned
            // address at or below it (inside the pixel's 96: lane offsets 6l), split by the
            // offset. Only the low two bytes of each word are used.
            const uintptr_t a = reinterpret_cast<uintptr_t>(q);
            const uint32_t off = static_cast<uint32_t>(a & 3u);  // 0 or 2
            const u32x2a v = *reinterpret_cast<const LUMEN_G u32x2a*>(q - off);
            const uint32_t lo = __builtin_amdgcn_alignbyte(v.y, v.x, off);
            c.w[3 * ch + 0] = lo;
            c.w[3 * ch + 1] = lo >> 16;
            c.w[3 * ch + 2] = v.y >> (8 * off);
        } else if constexpr (Q == 4) {
            c.w[ch] = *reinterpret_cast<const LUMEN_G uint32_t*>(q);
        } else if constexpr (Q == 8) {
            const u32x2 v = *reinterpret_cast<const LUMEN_G u32x2*>(q);
            c.w[2 * ch + 0] = v.x;
            c.w[2 * ch + 1] = v.y;
        } else {
            static_assert(Q == 16, "chunk of 2, 4, 8 or 16 disparities");
            const u32x4 v = *reinterpret_cast<const LUMEN_G u32x4*>(q);
            c.w[4 * ch + 0] = v.x;
            c.w[4 * ch + 1] = v.y;
            c.w[4 * ch + 2] = v.z;
            c.w[4 * ch + 3] = v.w;
        }
    }
    return c;
}

// p: the pixel's path-cost vector + Q*l; odd: odd_chunk_offset(l). Each chunk store writes a
// whole contiguous run of the scanline's lanes, so the stores are nontemporal.
template <int K, int LPS = 16>
__device__ __forceinline__ void store_path(LUMEN_G uint16_t* p, const u16x2 (&v)[K / 2], int odd = 0) {
    constexpr int Q = slice_q<K, LPS>();
#pragma unroll
    for (int ch = 0; ch < K / Q; ++ch) {
        LUMEN_G uint16_t* q = p + LPS * Q * ch + (snake_chunks<K, LPS>() && (ch & 1) ? odd : 0);
        const int b = ch * (Q / 2);
        if constexpr (Q == 2) {
            __builtin_nontemporal_store(bits(v[b]), reinterpret_cast<LUMEN_G uint32_t*>(q));
        } else if constexpr (Q == 6) {
            __builtin_nontemporal_store(u32x3{bits(v[b]), bits(v[b + 1]), bits(v[b + 2])},
                                        reinterpret_cast<LUMEN_G u32x3*>(q));
        } else if constexpr (Q == 4) {
            __builtin_nontemporal_store(u32x2{bits(v[b]), bits(v[b + 1])},
                                        reinterpret_cast<LUMEN_G u32x2*>(q));
        } else {
#pragma unroll
            for (int h = 0; h < Q / 8; ++h)
                __builtin_nontemporal_store(u32x4{bits(v[b + 4 * h]), bits(v[b + 4 * h + 1]),
                                                  bits(v[b + 4 * h + 2]), bits(v[b + 4 * h + 3])},
                                            reinterpret_cast<LUMEN_G u32x4*>(q + 8 * h));
        }
    }
}

// uint8 costs -> packed t16(C * costScale).
template <int K, int LPS = 16>
__device__ __forceinline__ void expand_cost(const CostBytes<K, LPS>& c, u16x2 cs,
                                            u16x2 (&out)[K / 2]) {
    if constexpr (cost_pairs<K, LPS>()) {
#pragma unroll
        for (int i = 0; i < K / 2; ++i)
            out[i] = as16(__builtin_amdgcn_perm(0u, c.w[i], 0x0c010c00u)) * cs;
    } else {
#pragma unroll
        for (int i = 0; i < K / 4; ++i) {
            out[2 * i] = as16(__builtin_amdgcn_perm(0u, c.w[i], 0x0c010c00u)) * cs;
            out[2 * i + 1] = as16(__builtin_amdgcn_perm(0u, c.w[i], 0x0c030c02u)) * cs;
        }
    }
}

// d-1 and d+1 neighbour vectors of a packed slice (lane l of its scanline's lanes). Inside a
// chunk they are the lane's own neighbouring elements. At a chunk's first element, d-1 is the
// last element of the same chunk in the lane before it in the chunk's direction; at a chunk's
// last element, d+1 is the first element of the same chunk in the lane after it. At a chunk's
// edge lanes the neighbour lies in the adjacent chunk: 16-lane rows (one row rotation per
// chunk each way) take it from the rotated value of that chunk (lane 0's d-1 is the last lane's
// last element of the previous chunk); 32- and 64-lane scanlines (whole-wave rotations by one
// lane, which cross the boundaries between the DPP rows of a scanline) alternate the chunks'
// direction (snake_chunks), so there the neighbour is the lane's own element of the adjacent
// chunk. The values the lowest lane gets for d = -1 and the highest for d = Dp (and any value
// rotated in from the other scanline of a 32-lane wave, which only reaches those lanes) are
// masked by the caller.
template <int K, int LPS = 16>
__device__ __forceinline__ void neighbours(const u16x2 (&v)[K / 2], uint32_t (&dm1)[K / 2],
                                           uint32_t (&dp1)[K / 2], int l) {
    constexpr int Q = slice_q<K, LPS>(), NC = K / Q, PQ = Q / 2;
    uint32_t fromPrev[NC], fromNext[NC];
    if constexpr (LPS >= 32) {
        static_assert(NC == 1 || snake_chunks<K, LPS>(), "multi-chunk wide slices alternate");
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const bool odd = (c & 1) != 0;
            const uint32_t lastW = bits(v[c * PQ + PQ - 1]), firstW = bits(v[c * PQ]);
            // Even chunks run up the lanes (d-1 from lane l-1), odd ones down (from lane l+1).
            uint32_t pv = odd ? dpp<kWaveRol1>(lastW) : dpp<kWaveRor1>(lastW);
            uint32_t nv = odd ? dpp<kWaveRor1>(firstW) : dpp<kWaveRol1>(firstW);
            const int lowLane = odd ? LPS - 1 : 0;   // holds the chunk's lowest disparities
            const int highLane = odd ? 0 : LPS - 1;  // holds its highest
            if (c > 0) pv = l == lowLane ? bits(v[(c - 1) * PQ + PQ - 1]) : pv;
            if (c < NC - 1) nv = l == highLane ? bits(v[(c + 1) * PQ]) : nv;
            fromPrev[c] = pv;
            fromNext[c] = nv;
        }
    } else {
        uint32_t prevLast[NC], nextFirst[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            prevLast[c] = dpp<kRowRor1>(bits(v[c * PQ + PQ - 1]));
            nextFirst[c] = dpp<kRowRor15>(bits(v[c * PQ]));
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            fromPrev[c] = (c > 0 && l == 0) ? prevLast[c - 1] : prevLast[c];
            fromNext[c] = (c < NC - 1 && l == LPS - 1) ? nextFirst[c + 1] : nextFirst[c];
        }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
#pragma unroll
        for (int jj = 0; jj < PQ; ++jj) {
            const int q = c * PQ + jj;
            const uint32_t cur = bits(v[q]);
            const uint32_t lo = jj == 0 ? fromPrev[c] : bits(v[q - 1]);
            const uint32_t hi = jj == PQ - 1 ? fromNext[c] : bits(v[q + 1]);
            dm1[q] = (cur << 16) | (lo >> 16);
            dp1[q] = (cur >> 16) | (hi << 16);
        }
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

// Minimum over the LPS lanes of a scanline, left in each of them: the DPP row minimum; for 32
// or 64 lanes the minimum with the other row of the pair (v_permlane16_swap exchanges rows 0
// and 1, 2 and 3); for 64 lanes then with the other pair (v_permlane32_swap exchanges the
// wave's halves).
template <int LPS>
__device__ __forceinline__ uint32_t group_min(uint32_t v) {
    v = row_min(v);
    if constexpr (LPS >= 32) {
        const auto sw = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        v = min(static_cast<uint32_t>(sw[0]), static_cast<uint32_t>(sw[1]));
    }
    if constexpr (LPS == 64) {
        const auto sw = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        v = min(static_cast<uint32_t>(sw[0]), static_cast<uint32_t>(sw[1]));
    }
    return v;
}

// The 256-entry t16(P2adj) table (src/sgbm_kernels.cu:304-310) lives in registers: lane l
// computes entries 4l..4l+3 once per wave into two packed words; a lookup fetches the word
// from lane g/4 with ds_bpermute.
struct P2Table {
    uint32_t lo, hi;
};

__device__ __forceinline__ P2Table make_p2(int P1, int P2, int lane) {
    uint32_t e[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int g = 4 * lane + k;
        const int v = max(P1 + 1, P2 * 8 / max(8, g));
        e[k] = static_cast<uint32_t>(v) & 0xFFFFu;
    }
    return P2Table{e[0] | (e[1] << 16), e[2] | (e[3] << 16)};
}

__device__ __forceinline__ uint32_t p2_lookup(const P2Table& t, uint32_t g) {
    const int src = static_cast<int>(g >> 2) * 4;  // byte address of the source lane
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(src, static_cast<int>(t.lo)));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(src, static_cast<int>(t.hi)));
    const uint32_t w = (g & 2u) ? hi : lo;
    return (g & 1u) ? (w >> 16) : (w & 0xFFFFu);
}

// Scanlines s0 .. s0 + 64/LPS - 1 of `path` of pair j (one per LPS lanes); wave-wide (no block
// synchronisation). Lanes whose scanline does not exist compute on a copy of pixel 0 and store
// to `trash`. PF is the prefetch depth and LPS the lanes per scanline (schedule choices; the
// results do not depend on them). HDIR = +1 or -1 compiles a horizontal path (LR or RL) with
// its step known: consecutive cost and path-cost vectors are then a constant LPS*K (= Dp)
// elements apart, so a block's loads and stores address off one pointer with immediate
// offsets instead of a 64-bit add each. HDIR = 0 serves every path with the step at run time.
template <int K, bool PAD, int PF = default_prefetch(K), int LPS = 16, int HDIR = 0>
__device__ void paths_wave(const PairJob& j, int path, int s0, uint16_t* trash_base) {
    static_assert(LPS == 16 || LPS == 32 || LPS == 64, "a scanline spans one, two or four DPP rows");
    static_assert(K % 2 == 0, "a lane holds whole packed pairs of disparities");
    static_assert(valid_prefetch(K, PF, LPS), "unsupported prefetch depth for this slice");
    static_assert(HDIR >= -1 && HDIR <= 1, "HDIR: 0 run-time step, +1 LR, -1 RL");
    const int lane = threadIdx.x & 63;
    const int g = lane / LPS, l = lane & (LPS - 1);
    const int s = s0 + g;
    int p0, stepv, len;
    scan_geom(path, s, j.W, j.H, p0, stepv, len);
    const int step = HDIR != 0 ? HDIR : stepv;
    int maxLen = max(len, __shfl_xor(len, 16));
    maxLen = max(maxLen, __shfl_xor(maxLen, 32));
    maxLen = __builtin_amdgcn_readfirstlane(maxLen);
    int minLen = min(len, __shfl_xor(len, 16));
    minLen = min(minLen, __shfl_xor(minLen, 32));
    minLen = __builtin_amdgcn_readfirstlane(minLen);
    const P2Table p2t = make_p2(j.P1, j.P2, lane);

    const int Dp = HDIR != 0 ? LPS * K : j.Dp, D = j.D;  // j.Dp == LPS * K
    constexpr int Q = slice_q<K, LPS>();
    constexpr int kLastChunkWord = (K / Q - 1) * (Q / 2);  // first packed word of the last chunk
    // The padding Dp - D is under 64, or under 32 for Dp = 96 (padded_range).
    static_assert(LPS * Q >= 64, "padding must fit in the last chunk");
    const LUMEN_G uint8_t* cbase = gmem(j.cost) + l * Q;
    const int odd = odd_chunk_offset<K, LPS>(l);  // to the lane's place in odd chunks
    LUMEN_G uint16_t* Lbase = gmem(j.L) + (size_t)path * j.pathStride + l * Q;
    const u16x2 cs = splat(static_cast<uint32_t>(j.costScale));
    const u16x2 P1p = splat(static_cast<uint32_t>(j.P1));
    const uint32_t m0 = l == 0 ? 0xFFFFu : 0u;  // slot d = 0 has no d-1 neighbour
    uint32_t c2m[K / 2], padm[K / 2];
#pragma unroll
    for (int q = 0; q < K / 2; ++q) {
        const int dlo = slice_d<K, LPS>(l, 2 * q);
        c2m[q] = (dlo == D - 1 ? 0xFFFFu : 0u) | (dlo + 1 == D - 1 ? 0xFFFF0000u : 0u);
        padm[q] = (dlo >= D ? 0xFFFFu : 0u) | (dlo + 1 >= D ? 0xFFFF0000u : 0u);
    }

    // Origin: L = t16(C * costScale).
    const long long cstep = (long long)step * Dp;  // bytes between consecutive cost vectors
    const LUMEN_G uint8_t* lp = cbase + (size_t)p0 * Dp;  // cost pointer of the last issued load
    int lt = 0;                                     // step index of the last issued load
    LUMEN_G uint16_t* sp = Lbase + (size_t)p0 * Dp;         // path-cost pointer of the step
    // Stores of finished scanlines land in the trash: one Dp-sized vector per scanline.
    LUMEN_G uint16_t* const trash = gmem(trash_base) + g * LPS * K + l * Q;
    u16x2 Lv[K / 2];
    {
        const CostBytes<K, LPS> c0 = load_cost<K, LPS>(lp, odd);
        expand_cost<K, LPS>(c0, cs, Lv);
        if constexpr (PAD) {
#pragma unroll
            for (int q = 0; q < K / 2; ++q) Lv[q] = as16(bits(Lv[q]) | padm[q]);
        }
        store_path<K, LPS>(len > 0 ? sp : trash, Lv, odd);
    }
    uint32_t M = group_min<LPS>(slice_min<K>(Lv));

    // Two register blocks of PF steps alternate: while one block is computed, the costs of the
    // next are in flight. Lane l of a scanline also fetches the two image bytes of step
    // tb + (l % PF) and looks up t16(P2adj) once per block; each step broadcasts its lane's
    // value with DPP row_newbcast (PF <= 16, so every row of a 32- or 64-lane scanline holds the
    // same values). CHECKED blocks handle scanline ends: loads past the end re-read the last pixel,
    // and steps past the end are computed but stored to the trash slot, so every store is
    // unconditional and the wait counts stay exact. While every scanline of the wave has all of
    // a block's loads and stores inside it, the blocks skip those clamps (no per-step compares
    // or 64-bit selects).
    const int last = len > 0 ? len - 1 : 0;
    const int lu = l & (PF - 1);
    auto fetch = [&](auto checked, int tb, CostBytes<K, LPS>(&dst)[PF], int (&iv)[2]) {
        if constexpr (decltype(checked)::value) {
            const int ti = min(tb + lu, last);
            iv[0] = gmem(j.left)[p0 + ti * step];
            iv[1] = gmem(j.left)[p0 + max(ti - 1, 0) * step];
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                const bool adv = lt < last;
                lp = adv ? lp + cstep : lp;
                lt += adv ? 1 : 0;
                dst[u] = load_cost<K, LPS>(lp, odd);
            }
        } else {
            const int ti = tb + lu;  // >= 1 and <= last
            iv[0] = gmem(j.left)[p0 + ti * step];
            iv[1] = gmem(j.left)[p0 + (ti - 1) * step];
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                lp += cstep;
                // With a run-time step, each load's address is the previous one plus the
                // step (one add): left to itself the compiler keeps base + u*step for every
                // u of the block in registers, 2*PF registers that cost waves per SIMD. The
                // empty asm only makes the pointer opaque; it emits nothing.
                if constexpr (HDIR == 0) asm volatile("" : "+v"(lp));
                dst[u] = load_cost<K, LPS>(lp, odd);
            }
            lt += PF;
        }
    };
    auto run = [&](auto checked, int tb, const CostBytes<K, LPS>(&blk)[PF], const int (&iv)[2]) {
        const uint32_t p2v = p2_lookup(p2t, static_cast<uint32_t>(abs(iv[0] - iv[1])));
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            sp += step * Dp;
            if constexpr (HDIR == 0) asm volatile("" : "+v"(sp));  // as for lp in fetch
            const u16x2 c3 = bcast_lo(M + row_bcast(p2v, u));
            const u16x2 Mp = bcast_lo(M);
            u16x2 Ch[K / 2];
            expand_cost<K, LPS>(blk[u], cs, Ch);
            // L(d -+ 1) + P1 are the neighbours of L + P1: one add per word instead of two.
            u16x2 Lp1[K / 2];
#pragma unroll
            for (int q = 0; q < K / 2; ++q) Lp1[q] = Lv[q] + P1p;
            uint32_t dm1[K / 2], dp1[K / 2];
            neighbours<K, LPS>(Lp1, dm1, dp1, l);
#pragma unroll
            for (int q = 0; q < K / 2; ++q) {
                uint32_t c1 = dm1[q];
                uint32_t c2 = dp1[q];
                if (q == 0) c1 |= m0;
                // Padding (d >= D) and the last disparity D - 1 lie in the last chunk (see
                // padded_range: a chunk spans more than Dp - D). The other words' masks are 0.
                if constexpr (PAD) {
                    if (q >= kLastChunkWord) c2 |= c2m[q];
                } else {
                    if (q == K / 2 - 1) c2 |= c2m[q];
                }
                const u16x2 mn = vmin(vmin(Lv[q], as16(c1)), vmin(as16(c2), c3));
                u16x2 nv = Ch[q] + mn - Mp;
                if constexpr (PAD) {
                    if (q >= kLastChunkWord) nv = as16(bits(nv) | padm[q]);
                }
                Lv[q] = nv;
            }
            if constexpr (decltype(checked)::value)
                store_path<K, LPS>(tb + u < len ? sp : trash, Lv, odd);
            else
                store_path<K, LPS>(sp, Lv, odd);
            M = group_min<LPS>(slice_min<K>(Lv));
        }
    };
    using Checked = std::integral_constant<bool, true>;
    using Unchecked = std::integral_constant<bool, false>;
    CostBytes<K, LPS> bufA[PF], bufB[PF];
    int gA[2], gB[2];
    fetch(Checked{}, 1, bufA, gA);
    // Entering the loop with block A complete keeps the loop-header state identical on both
    // edges, so no full drain is needed there.
    LUMEN_WAIT_VMCNT(0);
    // Block A's loads were issued before block B's stores: waiting until only those stores may
    // be outstanding retires the loads without waiting on the stores.
    constexpr int N = PF * path_stores(K, LPS);
    static_assert(N < 64, "vmcnt is 6 bits");
    int t0 = 1;
    // An iteration loads steps up to t0 + 3PF - 1 and stores steps up to t0 + 2PF - 1.
    for (; t0 + 3 * PF <= minLen; t0 += 2 * PF) {
        // The scheduling barriers keep each block's loads where they are issued: moved next to
        // their first use they would stall there.
        fetch(Unchecked{}, t0 + PF, bufB, gB);
        __builtin_amdgcn_sched_barrier(0);
        run(Unchecked{}, t0, bufA, gA);
        fetch(Unchecked{}, t0 + 2 * PF, bufA, gA);
        __builtin_amdgcn_sched_barrier(0);
        run(Unchecked{}, t0 + PF, bufB, gB);
        LUMEN_WAIT_VMCNT(N);
    }
    for (; t0 < maxLen; t0 += 2 * PF) {
        fetch(Checked{}, t0 + PF, bufB, gB);
        __builtin_amdgcn_sched_barrier(0);
        run(Checked{}, t0, bufA, gA);
        fetch(Checked{}, t0 + 2 * PF, bufA, gA);
        __builtin_amdgcn_sched_barrier(0);
        run(Checked{}, t0 + PF, bufB, gB);
        LUMEN_WAIT_VMCNT(N);
    }
}

// Issue priority of a wave by the length of its longest scanline: a launch ends when its
// longest scanlines end, so those waves go first within their SIMD.
__device__ __forceinline__ void set_priority_for_length(int len) {
    if (len >= 1536)
        __builtin_amdgcn_s_setprio(3);
    else if (len >= 768)
        __builtin_amdgcn_s_setprio(2);
    else if (len >= 256)
        __builtin_amdgcn_s_setprio(1);
    else
        __builtin_amdgcn_s_setprio(0);
}

// ---------------------------------------------------------------------------------------
// Path sum + left WTA + right WTA + LR check of one image row, block-wide.
//
// Each DPP row of the block takes one pixel at a time: it sums the path costs (mod 2^16),
// finds the left winner with a 4-step butterfly over (cost<<16 | d) keys plus the second
// minimum (src/sgbm_kernels.cu:748-805), and offers every (pixel, d) to right pixel
// x - m - d as the 64-bit key cost<<48 | d<<32 | S(d-1)<<16 | S(d+1) through an LDS atomic
// min. Smallest key = smallest cost, then smallest d, which is the first-minimum rule of the
// right WTA (src/sgbm_kernels.cu:820-834); keys of cost 65535 never count, as they never beat
// the reference's initial value. The LR check (src/sgbm_kernels.cu:862-890)
// only reads the same row, so it runs here too.
// ---------------------------------------------------------------------------------------
// LDS slot of right pixel xr in the right-WTA table. The lanes of a DPP row offer keys to
// right pixels Q apart; with Q = 4 or 8 one padding slot every Q keeps them on distinct banks,
// with Q = 6 (48-byte stride: 12 dwords, and 12 l mod 64 differs for the 16 lanes) they
// already are.
template <int Q>
__host__ __device__ constexpr int wta_slot(int xr) {
    return Q == 6 ? xr : xr + xr / Q;
}

template <int K>
__host__ __device__ inline size_t wta_lds_bytes(int W) {
    return (size_t)(wta_slot<slice_q<K>()>(W) + 1) * 8 + (size_t)W * 4;
}

template <int K>
__device__ __forceinline__ uint32_t pick32(const uint32_t (&v)[K / 2], int slot) {
    uint32_t r = 0xFFFFu;
#pragma unroll
    for (int q = 0; q < K / 2; ++q) {
        r = slot == 2 * q ? (v[q] & 0xFFFFu) : r;
        r = slot == 2 * q + 1 ? (v[q] >> 16) : r;
    }
    return r;
}

// One pixel of a row for the DPP row of lane l: sums the NR path-cost vectors in `raw`
// (mod 2^16), finds the left winner and second minimum, writes the left maps, and offers every
// (pixel, d) to its right pixel in `rbest`. DUMP (the parity hook's instantiation) also stores
// the sum vector to sumOut ([H][W][Dp], the cost volume's layout), so the sum the WTA works on
// is the one compared.
template <int K, bool PAD, int NR, bool DUMP = false>
__device__ __forceinline__ void wta_pixel(const PairJob& j, int x, int l, size_t rowPix,
                                          const uint32_t (&raw)[NR][K / 2],
                                          const uint32_t (&padm)[K / 2], int16_t* sdl,
                                          unsigned long long* rbest, uint16_t* sumOut) {
    constexpr int Q = slice_q<K>();
    const int D = j.D, m = j.m;
    u16x2 S[K / 2];
#pragma unroll
    for (int q = 0; q < K / 2; ++q) {
        S[q] = as16(raw[0][q]);
#pragma unroll
        for (int r = 1; r < NR; ++r) S[q] += as16(raw[r][q]);
    }
    if constexpr (DUMP) store_path<K>(gmem(sumOut) + (rowPix + x) * j.Dp + l * Q, S);
    if constexpr (PAD) {
        // Padding lies in the last chunk, which spans more than Dp - D (padded_range).
        constexpr int kLastChunkWord = (K / Q - 1) * (Q / 2);
#pragma unroll
        for (int q = kLastChunkWord; q < K / 2; ++q) S[q] = as16(bits(S[q]) | padm[q]);
    }
    uint32_t sm1[K / 2], sp1[K / 2];
    neighbours<K>(S, sm1, sp1, l);

    // Left WTA: (first minimum, second minimum of the multiset). The first minimum is the
    // smallest (cost << 16 | d) key; the two smallest costs of the lane's multiset are tracked
    // on packed pairs (even and odd elements in the two halves, three packed operations per
    // word), then the halves are merged. keys[] are also the right WTA's upper key words.
    uint32_t keys[K];
    uint32_t k1 = 0xFFFFFFFFu;
    u16x2 a2 = as16(0xFFFFFFFFu), b2 = as16(0xFFFFFFFFu);
#pragma unroll
    for (int q = 0; q < K / 2; ++q) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t hiHalf = h ? (bits(S[q]) & 0xFFFF0000u) : (bits(S[q]) << 16);
            keys[2 * q + h] = hiHalf | static_cast<uint32_t>(slice_d<K>(l, 2 * q + h));
            k1 = min(k1, keys[2 * q + h]);
        }
        b2 = vmin(b2, __builtin_elementwise_max(a2, S[q]));
        a2 = vmin(a2, S[q]);
    }
    const uint32_t aLo = bits(a2) & 0xFFFFu, aHi = bits(a2) >> 16;
    uint32_t m2 = min(max(aLo, aHi), min(bits(b2) & 0xFFFFu, bits(b2) >> 16));
#define LUMEN_WTA_COMBINE(CTRL)                   \
    {                                             \
        const uint32_t ok = dpp<CTRL>(k1);        \
        const uint32_t om = dpp<CTRL>(m2);        \
        m2 = min(min(m2, om), max(k1, ok) >> 16); \
        k1 = min(k1, ok);                         \
    }
    LUMEN_WTA_COMBINE(kQuadXor1)
    LUMEN_WTA_COMBINE(kQuadXor2)
    LUMEN_WTA_COMBINE(kRowHalfMirror)
    LUMEN_WTA_COMBINE(kRowMirror)
#undef LUMEN_WTA_COMBINE
    const int bestD = static_cast<int>(k1 & 0xFFFFu);
    const uint32_t best = k1 >> 16;
    if (l == slice_lane<K>(bestD)) {
        int16_t out;
        const uint32_t ratio = static_cast<uint32_t>(j.uniquenessRatio);
        if (j.uniquenessRatio > 0 && m2 != 0xFFFFu && m2 <= best + (best * ratio) / 100u) {
            out = -1;
        } else {
            float sub = static_cast<float>(bestD);
            if (bestD > 0 && bestD < D - 1) {
                const int slot = slice_elem<K>(bestD);
                sub = parabola(sub, pick32<K>(sm1, slot), best, pick32<K>(sp1, slot));
            }
            out = f2i16((sub + static_cast<float>(m)) * 16.0f);
        }
        gmem(j.dispL)[rowPix + x] = out;
        gmem(j.wtaInt)[rowPix + x] = static_cast<int16_t>(bestD);
        sdl[x] = out;
    }

    // Right WTA candidates, without branches: every (x, d) is offered. Keys whose cost is
    // 65535 (padding d >= D, or a real cost of 65535, which the reference's strict minimum
    // never takes) are ignored when the row is decoded, so they need no test here; an
    // (x, d) whose right pixel x - m - d is left of the image is offered to pixel 0 with an
    // upper word of all ones, which no decoded key can lose to.
#pragma unroll
    for (int q = 0; q < K / 2; ++q) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int xr = x - m - slice_d<K>(l, 2 * q + h);
            const uint32_t khi = xr >= 0 ? keys[2 * q + h] : 0xFFFFFFFFu;
            // S(d-1) << 16 | S(d+1): half h of sm1 and of sp1.
            const uint32_t klo = h ? __builtin_amdgcn_perm(sm1[q], sp1[q], 0x07060302u)
                                   : __builtin_amdgcn_perm(sm1[q], sp1[q], 0x05040100u);
            const int xc = max(xr, 0);
            atomicMin(&rbest[wta_slot<Q>(xc)],
                      (static_cast<unsigned long long>(khi) << 32) | klo);
        }
    }
}

// The NP path-cost vectors of one pixel for lane l (chunk by chunk, see the slice layout).
template <int K, int NP>
__device__ __forceinline__ void wta_load(const PairJob& j, size_t pix, int l,
                                         uint32_t (&raw)[NP][K / 2]) {
    constexpr int Q = slice_q<K>();
    const LUMEN_G uint16_t* src = gmem(j.L) + pix * j.Dp + l * Q;
#pragma unroll
    for (int r = 0; r < NP; ++r) {
        const LUMEN_G uint16_t* p = src + (size_t)r * j.pathStride;
#pragma unroll
        for (int ch = 0; ch < K / Q; ++ch) {
            if constexpr (Q == 8) {
                const u32x4 v =
                    __builtin_nontemporal_load(reinterpret_cast<const LUMEN_G u32x4*>(p + 16 * Q * ch));
                raw[r][4 * ch + 0] = v.x;
                raw[r][4 * ch + 1] = v.y;
                raw[r][4 * ch + 2] = v.z;
                raw[r][4 * ch + 3] = v.w;
            } else if constexpr (Q == 6) {
                const u32x3 v =
                    __builtin_nontemporal_load(reinterpret_cast<const LUMEN_G u32x3*>(p + 16 * Q * ch));
                raw[r][3 * ch + 0] = v.x;
                raw[r][3 * ch + 1] = v.y;
                raw[r][3 * ch + 2] = v.z;
            } else {
                const u32x2 v =
                    __builtin_nontemporal_load(reinterpret_cast<const LUMEN_G u32x2*>(p + 16 * Q * ch));
                raw[r][2 * ch + 0] = v.x;
                raw[r][2 * ch + 1] = v.y;
            }
        }
    }
}

// NP is the path count (4 or 8; the batched kernel switches on the pair's count): a
// compile-time count puts every plane's load of a pixel in flight before the first add, and
// while a pixel's path costs fit 32 registers the pixel loop is software-pipelined (the next
// pixel's loads are issued before the current pixel is processed).
template <int K, bool PAD, int NP, bool DUMP = false>
__device__ void wta_row(const PairJob& j, int y, unsigned long long* rbest,
                        uint16_t* sumOut = nullptr) {
    static_assert(NP == 4 || NP == 8, "the path count is 4 or 8");
    // rbest[xr + xr/Q]: running right-WTA key of right pixel xr. One padding entry every Q
    // entries spreads the 16 lanes of a row (whose candidates are Q right pixels apart) over
    // distinct LDS banks.
    constexpr int Q = slice_q<K>();
    const int W = j.W, D = j.D, m = j.m;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nslots = wta_slot<Q>(W) + 1;
    int16_t* sdl = reinterpret_cast<int16_t*>(rbest + nslots);
    int16_t* sdr = sdl + W;
    for (int i = tid; i < nslots; i += nt) rbest[i] = ~0ull;
    __syncthreads();
    const int g = tid >> 4, l = tid & 15, ngroups = nt >> 4;
    uint32_t padm[K / 2];
#pragma unroll
    for (int q = 0; q < K / 2; ++q) {
        const int dlo = slice_d<K>(l, 2 * q);
        padm[q] = (dlo >= D ? 0xFFFFu : 0u) | (dlo + 1 >= D ? 0xFFFF0000u : 0u);
    }
    const size_t rowPix = (size_t)y * W;
    if constexpr (NP * K / 2 <= 32) {
        uint32_t bufA[NP][K / 2], bufB[NP][K / 2];
        int x = g;
        if (x < W) wta_load<K, NP>(j, rowPix + x, l, bufA);
        for (; x < W; x += 2 * ngroups) {
            const int x1 = x + ngroups, x2 = x1 + ngroups;
            if (x1 < W) wta_load<K, NP>(j, rowPix + x1, l, bufB);
            wta_pixel<K, PAD, NP, DUMP>(j, x, l, rowPix, bufA, padm, sdl, rbest, sumOut);
            if (x2 < W) wta_load<K, NP>(j, rowPix + x2, l, bufA);
            if (x1 < W) wta_pixel<K, PAD, NP, DUMP>(j, x1, l, rowPix, bufB, padm, sdl, rbest, sumOut);
        }
    } else {
        // Wide slices: a second buffer would cost more occupancy than the overlap gains.
        uint32_t buf[NP][K / 2];
        for (int x = g; x < W; x += ngroups) {
            wta_load<K, NP>(j, rowPix + x, l, buf);
            wta_pixel<K, PAD, NP, DUMP>(j, x, l, rowPix, buf, padm, sdl, rbest, sumOut);
        }
    }
    __syncthreads();
    for (int xr = tid; xr < W; xr += nt) {
        const unsigned long long key = rbest[wta_slot<Q>(xr)];
        float sub = 0.0f;
        if ((key >> 48) != 0xFFFFu) {  // a candidate of cost below 65535
            const int bd = static_cast<int>((key >> 32) & 0xFFFFu);
            sub = static_cast<float>(bd);
            if (bd > 0 && bd < D - 1)
                sub = parabola(sub, static_cast<uint32_t>((key >> 16) & 0xFFFFu),
                               static_cast<uint32_t>(key >> 48), static_cast<uint32_t>(key & 0xFFFFu));
        }
        const int16_t dr = f2i16((sub + static_cast<float>(m)) * 16.0f);
        gmem(j.dispR)[rowPix + xr] = dr;
        sdr[xr] = dr;
    }
    __syncthreads();
    for (int x = tid; x < W; x += nt) {
        int16_t dl = sdl[x];
        if (j.doLR && dl > 0) {
            const int xr = x - dl / 16;
            if (xr < 0 || xr >= W) {
                dl = -1;
            } else {
                const int16_t dr = sdr[xr];
                if (dr <= 0 || abs(int(dl) - int(dr)) > j.lrMaxDiff) dl = -1;
            }
        }
        gmem(j.dispLR)[rowPix + x] = dl;
    }
}

// ---------------------------------------------------------------------------------------
// Speckle filter (src/sgbm_kernels.cu:979-1013) with snapshot semantics: reads `in`, writes
// `out`, so the result does not depend on thread timing. A pixel with a value > 0 keeps it when
// at least minSimilar pixels of its 5x5 window (itself included, out-of-image ones skipped)
// are > 0 and within maxDiff of it, and becomes -1 otherwise.
// ---------------------------------------------------------------------------------------
// One TW x TH tile, block-wide: the tile and a 2-pixel halo of `in` are staged in LDS (`tile`,
// (TH+4) x (TW+4) entries) with one load per entry, then every pixel counts its 5x5 window
// there. Out-of-image neighbours are staged as 0, which the count skips (only values > 0
// count). Every thread of the block must call it with the same tile; it ends with a barrier,
// so `tile` may be reused at once.
template <int TW, int TH>
__device__ void speckle_tile(const LUMEN_G int16_t* __restrict__ in, LUMEN_G int16_t* __restrict__ out,
                             int W, int H, int x0, int y0, int minSimilar, int maxDiff,
                             int16_t* tile) {
    constexpr int PW = TW + 4, PH = TH + 4;
    for (int i = threadIdx.x; i < PW * PH; i += blockDim.x) {
        const int ty = i / PW, tx = i - ty * PW;
        const int gx = x0 + tx - 2, gy = y0 + ty - 2;
        tile[i] = (gx >= 0 && gx < W && gy >= 0 && gy < H) ? in[(size_t)gy * W + gx] : int16_t(0);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < TW * TH; i += blockDim.x) {
        const int ty = i / TW, tx = i - ty * TW;
        const int x = x0 + tx, y = y0 + ty;
        if (x >= W || y >= H) continue;
        const int16_t* w = tile + ty * PW + tx;  // top-left of the 5x5 window
        const int16_t c = w[2 * PW + 2];
        int16_t res = c;
        if (c > 0) {
            int n = 0;
#pragma unroll
            for (int dy = 0; dy < 5; ++dy)
#pragma unroll
                for (int dx = 0; dx < 5; ++dx) {
                    const int16_t v = w[dy * PW + dx];
                    n += (v > 0 && abs(int(c) - int(v)) <= maxDiff) ? 1 : 0;
                }
            if (n < minSimilar) res = -1;
        }
        out[(size_t)y * W + x] = res;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------------------
// 3x3 median of positive values (src/sgbm_kernels.cu:892-935) over one 64x4 tile,
// block-wide (256 threads). The tile and a 1-pixel halo are staged in LDS; out-of-image taps
// are staged as 0, which the positive-only rule excludes. The nine taps (non-positive ones
// replaced by INT_MAX so they sort last) go through a fixed 25-comparator sorting network;
// the upper median is element n/2 of the sorted positives.
// ---------------------------------------------------------------------------------------
constexpr int kMedTW = 64, kMedTH = 4;

__device__ __forceinline__ void cswap(int& a, int& b) {
    const int lo = min(a, b);
    b = max(a, b);
    a = lo;
}

// `mirror`, when not null, receives a second copy of the output.
__device__ void median_tile(const LUMEN_G int16_t* __restrict__ in, LUMEN_G int16_t* __restrict__ out, int W,
                            int H, int x0, int y0, int16_t (*tile)[kMedTW + 2],
                            LUMEN_G int16_t* __restrict__ mirror = nullptr) {
    for (int i = threadIdx.x; i < (kMedTH + 2) * (kMedTW + 2); i += blockDim.x) {
        const int ty = i / (kMedTW + 2), tx = i % (kMedTW + 2);
        const int gx = x0 + tx - 1, gy = y0 + ty - 1;
        tile[ty][tx] = (gx >= 0 && gx < W && gy >= 0 && gy < H) ? in[(size_t)gy * W + gx] : int16_t(0);
    }
    __syncthreads();
    const int tx = threadIdx.x % kMedTW, ty = threadIdx.x / kMedTW;
    const int x = x0 + tx, y = y0 + ty;
    if (ty < kMedTH && x < W && y < H) {
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
        for (int q = 1; q < 9; ++q) r = k == q ? v[q] : r;
        const int16_t o = n == 0 ? tile[ty + 1][tx + 1] : static_cast<int16_t>(r);
        out[(size_t)y * W + x] = o;
        if (mirror) mirror[(size_t)y * W + x] = o;
    }
}

}  // namespace lumen
