// rectify.hip - the rectification remap (the stage before the matcher).
//
// Every output pixel reads its two map coordinates (8 bytes), gathers four source taps and
// writes its value, so the remap is bound by HBM at about 10 bytes per pixel and channel-1
// image: the taps come from a source image that stays in L2 (a 1080p image is 2 MB), the map
// and output streams do not. For grey images a thread takes four consecutive pixels (two
// 16-byte map loads and one 4-byte store); a rectifier launches both images of a pair in one
// grid (blockIdx.y selects the image).
#include "lumen_rectify.h"
#include "scene_common.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>

namespace lumen {
namespace {

constexpr int kBlock = 256;
constexpr int kPixPerThread = 4;

struct RemapImage {
    const uint8_t* src;
    const float* mapx;
    const float* mapy;
    uint8_t* dst;
};

struct RemapPair {
    RemapImage img[2];
};

// Taps of a map coordinate, or none when it is NaN or beyond +-2^30 (every tap outside the
// image; the reference's float -> int conversion is undefined there).
struct Taps {
    int x0, y0;
    float w00, w01, w10, w11;
    bool ok;
};

// src/rectification.cu:218-234: floor, fractional parts, the four weights as written.
__device__ __forceinline__ Taps taps_of(float sx, float sy) {
    Taps t;
    t.ok = sx >= -1073741824.0f && sx <= 1073741824.0f && sy >= -1073741824.0f &&
           sy <= 1073741824.0f;
    const float fx = t.ok ? floorf(sx) : 0.0f, fy = t.ok ? floorf(sy) : 0.0f;
    t.x0 = static_cast<int>(fx);
    t.y0 = static_cast<int>(fy);
    const float dx = sx - static_cast<float>(t.x0), dy = sy - static_cast<float>(t.y0);
    t.w00 = (1.0f - dx) * (1.0f - dy);
    t.w01 = (1.0f - dx) * dy;
    t.w10 = dx * (1.0f - dy);
    t.w11 = dx * dy;
    return t;
}

// src/rectification.cu:236-255: taps outside the image are 0; the weighted sum in the order
// written (no fused multiply-add: the build disables contraction), + 0.5, truncated.
__device__ __forceinline__ uint8_t sample(const uint8_t* __restrict__ src, int W, int H, int C,
                                          const Taps& t, int c) {
    if (!t.ok) return 0;
    const int x1 = t.x0 + 1, y1 = t.y0 + 1;
    const bool x0in = t.x0 >= 0 && t.x0 < W, x1in = x1 >= 0 && x1 < W;
    const bool y0in = t.y0 >= 0 && t.y0 < H, y1in = y1 >= 0 && y1 < H;
    const long long row0 = (long long)t.y0 * W * C, row1 = (long long)y1 * W * C;
    const float v00 = (x0in && y0in) ? static_cast<float>(src[row0 + (long long)t.x0 * C + c]) : 0.0f;
    const float v01 = (x0in && y1in) ? static_cast<float>(src[row1 + (long long)t.x0 * C + c]) : 0.0f;
    const float v10 = (x1in && y0in) ? static_cast<float>(src[row0 + (long long)x1 * C + c]) : 0.0f;
    const float v11 = (x1in && y1in) ? static_cast<float>(src[row1 + (long long)x1 * C + c]) : 0.0f;
    const float val = t.w00 * v00 + t.w01 * v01 + t.w10 * v10 + t.w11 * v11;
    return static_cast<uint8_t>(static_cast<unsigned>(val + 0.5f));
}

// Grey sample for a 4-byte aligned source of nbytes = W*H bytes. When all four taps are in the
// image, each row's pair of taps is read as the two dwords at the pair's 4-byte aligned address
// (the pair sits at byte offset 0-3 of those 8 bytes) and extracted with one byte-align: two
// loads per row instead of two byte gathers. Same bytes, same arithmetic as sample(); taps at
// the image's border, and a pair whose 8 bytes would end past the image, take sample().
__device__ __forceinline__ uint8_t sample_grey(const uint8_t* __restrict__ src, int W, int H,
                                               long long nbytes, const Taps& t) {
    if (t.ok && t.x0 >= 0 && t.x0 + 1 < W && t.y0 >= 0 && t.y0 + 1 < H) {
        const long long a0 = (long long)t.y0 * W + t.x0, a1 = a0 + W;
        const long long b0 = a0 & ~3LL, b1 = a1 & ~3LL;
        if (b1 + 8 <= nbytes) {
            const uint32_t* p0 = reinterpret_cast<const uint32_t*>(src + b0);
            const uint32_t* p1 = reinterpret_cast<const uint32_t*>(src + b1);
            const uint32_t r0 = __builtin_amdgcn_alignbyte(p0[1], p0[0], static_cast<uint32_t>(a0 - b0));
            const uint32_t r1 = __builtin_amdgcn_alignbyte(p1[1], p1[0], static_cast<uint32_t>(a1 - b1));
            const float v00 = static_cast<float>(r0 & 0xFFu), v10 = static_cast<float>((r0 >> 8) & 0xFFu);
            const float v01 = static_cast<float>(r1 & 0xFFu), v11 = static_cast<float>((r1 >> 8) & 0xFFu);
            const float val = t.w00 * v00 + t.w01 * v01 + t.w10 * v10 + t.w11 * v11;
            return static_cast<uint8_t>(static_cast<unsigned>(val + 0.5f));
        }
    }
    return sample(src, W, H, 1, t, 0);
}

// Grey images, four pixels per thread. VEC: maps 16-byte, output and source 4-byte aligned
// (checked on the host); the last partial group, and unaligned buffers, go pixel by pixel.
template <bool VEC>
__global__ void __launch_bounds__(kBlock) remap_grey_kernel(RemapPair p, int W, int H, long long n) {
    const RemapImage im = p.img[blockIdx.y];
    const long long i = ((long long)blockIdx.x * kBlock + threadIdx.x) * kPixPerThread;
    if (i >= n) return;
    if (VEC && i + kPixPerThread <= n) {
        const float4 mx = *reinterpret_cast<const float4*>(im.mapx + i);
        const float4 my = *reinterpret_cast<const float4*>(im.mapy + i);
        const uint32_t b0 = sample_grey(im.src, W, H, n, taps_of(mx.x, my.x));
        const uint32_t b1 = sample_grey(im.src, W, H, n, taps_of(mx.y, my.y));
        const uint32_t b2 = sample_grey(im.src, W, H, n, taps_of(mx.z, my.z));
        const uint32_t b3 = sample_grey(im.src, W, H, n, taps_of(mx.w, my.w));
        *reinterpret_cast<uint32_t*>(im.dst + i) = b0 | (b1 << 8) | (b2 << 16) | (b3 << 24);
        return;
    }
    for (int k = 0; k < kPixPerThread && i + k < n; ++k)
        im.dst[i + k] = sample(im.src, W, H, 1, taps_of(im.mapx[i + k], im.mapy[i + k]), 0);
}

// Interleaved C-channel images (C = 2..4, source 4-byte aligned): one pixel per thread. When
// all four taps are in the image, a row's two taps are the 2C bytes at x0*C, read as the three
// dwords at their 4-byte aligned address (the run starts at byte 0-3 and ends by byte 11) and
// split with two byte-aligns: six loads a pixel instead of 4C byte gathers. Same bytes and
// arithmetic as sample(); border taps, and a run whose 12 bytes would end past the image, take
// sample().
template <int C>
__global__ void __launch_bounds__(kBlock) remap_multi_fast_kernel(RemapPair p, int W, int H,
                                                                  long long n) {
    static_assert(C >= 2 && C <= 4, "two taps of C channels fit 8 bytes");
    const RemapImage im = p.img[blockIdx.y];
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const Taps t = taps_of(im.mapx[i], im.mapy[i]);
    const long long nbytes = n * C;
    if (t.ok && t.x0 >= 0 && t.x0 + 1 < W && t.y0 >= 0 && t.y0 + 1 < H) {
        const long long a0 = ((long long)t.y0 * W + t.x0) * C, a1 = a0 + (long long)W * C;
        const long long b0 = a0 & ~3LL, b1 = a1 & ~3LL;
        if (b1 + 12 <= nbytes) {
            // run(row): the 8 bytes starting at the row's first tap, as lo (bytes 0-3), hi (4-7).
            auto run = [&](long long b, long long a, uint32_t& lo, uint32_t& hi) {
                const uint32_t* q = reinterpret_cast<const uint32_t*>(im.src + b);
                const uint32_t sh = static_cast<uint32_t>(a - b);
                const uint32_t w0 = q[0], w1 = q[1], w2 = q[2];
                lo = __builtin_amdgcn_alignbyte(w1, w0, sh);
                hi = __builtin_amdgcn_alignbyte(w2, w1, sh);
            };
            uint32_t lo0, hi0, lo1, hi1;
            run(b0, a0, lo0, hi0);
            run(b1, a1, lo1, hi1);
            auto byte = [](uint32_t lo, uint32_t hi, int k) -> float {
                return static_cast<float>(((k < 4 ? lo : hi) >> (8 * (k & 3))) & 0xFFu);
            };
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const float v00 = byte(lo0, hi0, c), v10 = byte(lo0, hi0, C + c);
                const float v01 = byte(lo1, hi1, c), v11 = byte(lo1, hi1, C + c);
                const float val = t.w00 * v00 + t.w01 * v01 + t.w10 * v10 + t.w11 * v11;
                im.dst[i * C + c] = static_cast<uint8_t>(static_cast<unsigned>(val + 0.5f));
            }
            return;
        }
    }
    for (int c = 0; c < C; ++c) im.dst[i * C + c] = sample(im.src, W, H, C, t, c);
}

// Interleaved multi-channel images: one pixel per thread, every channel from the same taps.
__global__ void __launch_bounds__(kBlock) remap_multi_kernel(RemapPair p, int W, int H, int C,
                                                             long long n) {
    const RemapImage im = p.img[blockIdx.y];
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const Taps t = taps_of(im.mapx[i], im.mapy[i]);
    for (int c = 0; c < C; ++c) im.dst[i * C + c] = sample(im.src, W, H, C, t, c);
}

bool aligned(const void* p, size_t a) { return (reinterpret_cast<uintptr_t>(p) & (a - 1)) == 0; }

hipError_t launch_remap(const RemapPair& p, int nimg, int W, int H, int C, hipStream_t s) {
    const long long n = (long long)W * H;
    if (C == 1) {
        bool vec = true;
        for (int k = 0; k < nimg; ++k)
            vec = vec && aligned(p.img[k].mapx, 16) && aligned(p.img[k].mapy, 16) &&
                  aligned(p.img[k].dst, 4) && aligned(p.img[k].src, 4);
        const long long per = (long long)kBlock * kPixPerThread;
        const dim3 grid(static_cast<unsigned>((n + per - 1) / per), nimg);
        if (vec)
            remap_grey_kernel<true><<<grid, kBlock, 0, s>>>(p, W, H, n);
        else
            remap_grey_kernel<false><<<grid, kBlock, 0, s>>>(p, W, H, n);
    } else {
        bool src4 = true;
        for (int k = 0; k < nimg; ++k) src4 = src4 && aligned(p.img[k].src, 4);
        const dim3 grid(static_cast<unsigned>((n + kBlock - 1) / kBlock), nimg);
        if (src4 && C == 3)
            remap_multi_fast_kernel<3><<<grid, kBlock, 0, s>>>(p, W, H, n);
        else if (src4 && C == 4)
            remap_multi_fast_kernel<4><<<grid, kBlock, 0, s>>>(p, W, H, n);
        else if (src4 && C == 2)
            remap_multi_fast_kernel<2><<<grid, kBlock, 0, s>>>(p, W, H, n);
        else
            remap_multi_kernel<<<grid, kBlock, 0, s>>>(p, W, H, C, n);
    }
    return hipGetLastError();
}

// Image arguments of the remap entry points; the message is the one set on failure.
int check_image(const void* a, const void* b, int W, int H, int C) {
    if (!a || !b || W <= 0 || H <= 0) return (set_last_error("Input images are empty"), LUMEN_EINVAL);
    if (C < 1 || C > 4) return (set_last_error("Channel count must be 1 to 4"), LUMEN_EINVAL);
    if ((long long)W * H > (1LL << 31)) return (set_last_error("Image larger than 2^31 pixels"), LUMEN_EINVAL);
    return LUMEN_OK;
}

}  // namespace

SceneStaging& scene_staging() {
    static SceneStaging* st = new SceneStaging();  // never destroyed: see SceneStaging
    return *st;
}

}  // namespace lumen

using namespace lumen;

struct lumen_rectifier {
    int W = 0, H = 0;
    bool preRectified = true;
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {};
    float* maps[4] = {};          // left x, left y, right x, right y
    uint8_t* io[4] = {};          // host-form staging: left in, right in, left out, right out
    size_t ioCap = 0;

    int fail(int code, const std::string& msg) {
        err = msg;
        return code;
    }
    int hip_fail(hipError_t e, const char* what) {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? LUMEN_ENOMEM : LUMEN_EDEVICE;
    }
    ~lumen_rectifier() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto* m : maps)
            if (m) (void)hipFree(m);
        for (auto* b : io)
            if (b) (void)hipFree(b);
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
    int check_size(int width, int height, int channels) {
        if (!preRectified && (width != W || height != H))
            return fail(LUMEN_EINVAL, "Image size " + std::to_string(width) + "x" +
                                          std::to_string(height) +
                                          " does not match the rectification maps (" +
                                          std::to_string(W) + "x" + std::to_string(H) + ")");
        if (channels < 1 || channels > 4) return fail(LUMEN_EINVAL, "Channel count must be 1 to 4");
        return LUMEN_OK;
    }
    // Remaps (or, pre-rectified, copies) device images on the handle's stream.
    int run(const uint8_t* l, const uint8_t* r, int width, int height, int C, uint8_t* lo,
            uint8_t* ro) {
        const size_t bytes = (size_t)width * height * C;
        hipError_t e = hipEventRecord(ev[0], stream);
        if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
        if (preRectified) {
            e = hipMemcpyAsync(lo, l, bytes, hipMemcpyDeviceToDevice, stream);
            if (e == hipSuccess) e = hipMemcpyAsync(ro, r, bytes, hipMemcpyDeviceToDevice, stream);
            if (e != hipSuccess) return hip_fail(e, "copy");
        } else {
            RemapPair p{{{l, maps[0], maps[1], lo}, {r, maps[2], maps[3], ro}}};
            e = launch_remap(p, 2, W, H, C, stream);
            if (e != hipSuccess) return hip_fail(e, "remap kernel launch");
        }
        e = hipEventRecord(ev[1], stream);
        return e == hipSuccess ? LUMEN_OK : hip_fail(e, "hipEventRecord");
    }
};

extern "C" int lumen_rectifier_create(const float* map0x, const float* map0y, const float* map1x,
                                      const float* map1y, int width, int height,
                                      lumen_rectifier** out) {
    if (!out) return (set_last_error("out is NULL"), LUMEN_EINVAL);
    *out = nullptr;
    const bool pre = !map0x || !map0y;
    if (!pre && (!map1x || !map1y))
        return (set_last_error("Right camera maps are missing"), LUMEN_EINVAL);
    if (!pre && (width <= 0 || height <= 0))
        return (set_last_error("Invalid image size"), LUMEN_EINVAL);
    if (!pre && (long long)width * height > (1LL << 31))
        return (set_last_error("Image larger than 2^31 pixels"), LUMEN_EINVAL);
    auto* r = new (std::nothrow) lumen_rectifier();
    if (!r) return (set_last_error("out of host memory"), LUMEN_ENOMEM);
    r->W = width;
    r->H = height;
    r->preRectified = pre;
    hipError_t e = hipGetDevice(&r->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) e = hipEventCreate(&r->ev[k]);
    const float* host[4] = {map0x, map0y, map1x, map1y};
    const size_t bytes = (size_t)width * height * sizeof(float);
    for (int k = 0; k < 4 && e == hipSuccess && !pre; ++k) {
        e = hipMalloc(&r->maps[k], bytes);
        if (e == hipSuccess) e = hipMemcpy(r->maps[k], host[k], bytes, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        const int rc = scene_hip_fail(e, "rectifier set-up");
        delete r;
        return rc;
    }
    *out = r;
    return LUMEN_OK;
}

extern "C" int lumen_rectifier_destroy(lumen_rectifier* r) {
    delete r;
    return LUMEN_OK;
}

extern "C" int lumen_rectifier_is_pre_rectified(const lumen_rectifier* r) {
    return r ? (r->preRectified ? 1 : 0) : -1;
}

extern "C" int lumen_rectifier_rectify(lumen_rectifier* r, const uint8_t* left,
                                       const uint8_t* right, int width, int height, int channels,
                                       uint8_t* left_out, uint8_t* right_out) {
    if (!r) return (set_last_error("handle is NULL"), LUMEN_EINVAL);
    if (!left || !right || !left_out || !right_out || width <= 0 || height <= 0)
        return r->fail(LUMEN_EINVAL, "Input images are empty");
    if (int rc = r->check_size(width, height, channels)) return rc;
    const size_t bytes = (size_t)width * height * channels;
    if (bytes > r->ioCap) {
        for (auto*& b : r->io) {
            if (b) (void)hipFree(b);
            b = nullptr;
        }
        r->ioCap = 0;
        for (auto*& b : r->io) {
            hipError_t e = hipMalloc(&b, bytes);
            if (e != hipSuccess) {
                b = nullptr;
                return r->hip_fail(e, "rectifier buffers");
            }
        }
        r->ioCap = bytes;
    }
    hipError_t e = hipMemcpyAsync(r->io[0], left, bytes, hipMemcpyHostToDevice, r->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(r->io[1], right, bytes, hipMemcpyHostToDevice, r->stream);
    if (e != hipSuccess) return r->hip_fail(e, "upload");
    if (int rc = r->run(r->io[0], r->io[1], width, height, channels, r->io[2], r->io[3])) return rc;
    e = hipMemcpyAsync(left_out, r->io[2], bytes, hipMemcpyDeviceToHost, r->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(right_out, r->io[3], bytes, hipMemcpyDeviceToHost, r->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    return e == hipSuccess ? LUMEN_OK : r->hip_fail(e, "download");
}

extern "C" int lumen_rectifier_rectify_device(lumen_rectifier* r, const uint8_t* d_left,
                                              const uint8_t* d_right, int width, int height,
                                              int channels, uint8_t* d_left_out,
                                              uint8_t* d_right_out, int sync) {
    if (!r) return (set_last_error("handle is NULL"), LUMEN_EINVAL);
    if (!d_left || !d_right || !d_left_out || !d_right_out || width <= 0 || height <= 0)
        return r->fail(LUMEN_EINVAL, "Input images are empty");
    if (int rc = r->check_size(width, height, channels)) return rc;
    if (int rc = r->run(d_left, d_right, width, height, channels, d_left_out, d_right_out))
        return rc;
    if (sync) {
        hipError_t e = hipStreamSynchronize(r->stream);
        if (e != hipSuccess) return r->hip_fail(e, "hipStreamSynchronize");
    }
    return LUMEN_OK;
}

extern "C" int lumen_rectifier_get_stream(lumen_rectifier* r, void** stream_out) {
    if (!r || !stream_out) return (set_last_error("NULL argument"), LUMEN_EINVAL);
    *stream_out = r->stream;
    return LUMEN_OK;
}

extern "C" int lumen_rectifier_get_time_ms(lumen_rectifier* r, float* ms_out) {
    if (!r || !ms_out) return (set_last_error("NULL argument"), LUMEN_EINVAL);
    hipError_t e = hipEventSynchronize(r->ev[1]);
    if (e == hipSuccess) e = hipEventElapsedTime(ms_out, r->ev[0], r->ev[1]);
    return e == hipSuccess ? LUMEN_OK : r->hip_fail(e, "hipEventElapsedTime");
}

extern "C" const char* lumen_rectifier_last_error(const lumen_rectifier* r) {
    return r ? r->err.c_str() : "";
}

extern "C" int lumen_remap(const uint8_t* src, const float* mapx, const float* mapy, uint8_t* dst,
                           int width, int height, int channels) {
    if (int rc = check_image(src, dst, width, height, channels)) return rc;
    if (!mapx || !mapy) return (set_last_error("Remap tables are empty"), LUMEN_EINVAL);
    const size_t n = (size_t)width * height, bytes = n * channels;
    SceneStaging& st = scene_staging();
    std::lock_guard<std::mutex> lock(st.mu);
    hipError_t e = st.ensure(0, bytes);
    if (e == hipSuccess) e = st.ensure(1, n * 4);
    if (e == hipSuccess) e = st.ensure(2, n * 4);
    if (e == hipSuccess) e = st.ensure(3, bytes);
    if (e != hipSuccess) return scene_hip_fail(e, "device buffers");
    auto* ds = static_cast<uint8_t*>(st.buf[0]);
    auto* mx = static_cast<float*>(st.buf[1]);
    auto* my = static_cast<float*>(st.buf[2]);
    auto* dd = static_cast<uint8_t*>(st.buf[3]);
    e = hipMemcpyAsync(ds, src, bytes, hipMemcpyHostToDevice, st.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(mx, mapx, n * 4, hipMemcpyHostToDevice, st.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(my, mapy, n * 4, hipMemcpyHostToDevice, st.stream);
    if (e == hipSuccess) {
        RemapPair p{{{ds, mx, my, dd}, {}}};
        e = launch_remap(p, 1, width, height, channels, st.stream);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(dst, dd, bytes, hipMemcpyDeviceToHost, st.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st.stream);
    return e == hipSuccess ? LUMEN_OK : scene_hip_fail(e, "remap");
}

extern "C" int lumen_remap_device(const uint8_t* d_src, const float* d_mapx, const float* d_mapy,
                                  uint8_t* d_dst, int width, int height, int channels,
                                  void* stream) {
    if (int rc = check_image(d_src, d_dst, width, height, channels)) return rc;
    if (!d_mapx || !d_mapy) return (set_last_error("Remap tables are empty"), LUMEN_EINVAL);
    RemapPair p{{{d_src, d_mapx, d_mapy, d_dst}, {}}};
    hipError_t e = launch_remap(p, 1, width, height, channels, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? LUMEN_OK : scene_hip_fail(e, "remap kernel launch");
}
