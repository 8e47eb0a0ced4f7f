// sgbm_capi.hip - C ABI of the matcher: validation, device buffers, stage order, timing.
//
// Stage order follows StereoSGBM::compute (src/sgbm_gpu.cu:147-195):
//   upload -> cost (census or prefilter+SAD) -> all paths -> sum+WTA(L,R)+LR -> speckle ->
//   median -> download.
// Buffers grow on demand and are kept between calls (src/sgbm_gpu.cu:119-145 resizes lazily
// too). Nothing is zeroed per call: every buffer a stage reads was fully written by the stage
// before it.
#include "host_copy.h"
#include "lumen_sgbm.h"
#include "sgbm_host.h"
#include "sgbm_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace lumen;

namespace {

thread_local std::string g_last_error;

struct DevBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap && ptr) return hipSuccess;
        if (ptr) {
            hipError_t e = hipFree(ptr);
            ptr = nullptr;
            cap = 0;
            if (e != hipSuccess) return e;
        }
        if (bytes == 0) return hipSuccess;
        hipError_t e = hipMalloc(&ptr, bytes);
        if (e != hipSuccess) {
            ptr = nullptr;
            return e;
        }
        cap = bytes;
        return hipSuccess;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(ptr); }
};

// Pinned host staging: a pageable strided copy goes through the runtime's own staging row by
// row, which costs more than the whole frame's compute at 1080p.
struct PinnedBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap && ptr) return hipSuccess;
        if (ptr) {
            hipError_t e = hipHostFree(ptr);
            ptr = nullptr;
            cap = 0;
            if (e != hipSuccess) return e;
        }
        hipError_t e = hipHostMalloc(&ptr, bytes, hipHostMallocDefault);
        if (e != hipSuccess) {
            ptr = nullptr;
            return e;
        }
        cap = bytes;
        return hipSuccess;
    }
    void release() {
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(ptr); }
};

// Per compute: 0 before the upload, then each stage's first kernel start and last kernel end:
// 1-2 cost, 3-4 aggregation, 5-6 sum + WTA + LR check, 7-8 post-processing.
constexpr int kNumEvents = 9;
constexpr int kRing = 64;

struct Geometry {
    lumen_stereo_params params{};  // of the last prepared compute (the dump hook's frame)
    int W = 0, H = 0, D = 0, Dp = 0, m = 0, npaths = 0, costScale = 0;
    size_t pixels() const { return (size_t)W * H; }
    size_t volume() const { return pixels() * Dp; }
};

}  // namespace

struct lumen_sgbm {
    lumen_stereo_params params{};
    std::string err;
    bool ready = false;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[kRing][kNumEvents] = {};
    bool slotHasUpload[kRing] = {};  // the slot's compute uploaded host images (event 0 used)
    int slotTiming[kRing] = {};      // the timing level the slot's compute recorded
    int timing = LUMEN_TIMING_STAGES;
    int ringNext = 0;   // next slot to record
    int ringCount = 0;  // completed-or-pending calls recorded
    Geometry geo;
    bool haveResult = false;
    DevBuf left, right, cost, paths, dispL, dispR, dispLR, wtaInt, spk, out,
        sum, trash, dumpTmp;
    PinnedBuf stageIn, stageOut;
    bool hasSchedule = false;  // lumen_sgbm_set_schedule override
    Schedule schedule{};

    int fail(int code, const std::string& msg) {
        err = msg;
        return code;
    }
    int hip_fail(hipError_t e, const char* what) {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? LUMEN_ENOMEM : LUMEN_EDEVICE;
    }
    int ensure_ready() {
        if (ready) return LUMEN_OK;
        hipError_t e = hipGetDevice(&device);
        if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
        e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
        if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
        for (int i = 0; i < kRing; ++i)
            for (int j = 0; j < kNumEvents; ++j) {
                e = hipEventCreate(&ev[i][j]);
                if (e != hipSuccess) return hip_fail(e, "hipEventCreate");
            }
        ready = true;
        return LUMEN_OK;
    }
    ~lumen_sgbm() {
        if (ready) {
            (void)hipStreamSynchronize(stream);
            for (int i = 0; i < kRing; ++i)
                for (int j = 0; j < kNumEvents; ++j)
                    if (ev[i][j]) (void)hipEventDestroy(ev[i][j]);
            (void)hipStreamDestroy(stream);
        }
        DevBuf* all[] = {&left, &right, &cost, &paths, &dispL,   &dispR, &dispLR,
                         &wtaInt, &spk, &out, &sum, &trash, &dumpTmp};
        for (DevBuf* b : all) b->release();
        stageIn.release();
        stageOut.release();
    }
};

namespace {

// Validation of src/sgbm_gpu.cu:148-166, plus the limits of this implementation.
int validate(lumen_sgbm* h, const void* left, const void* right, int W, int H) {
    const std::string msg = validate_pair(h->params, W, H, left != nullptr && right != nullptr);
    if (!msg.empty()) return h->fail(LUMEN_EINVAL, msg);
    if (h->hasSchedule) {
        std::string why;
        const int K = padded_range(h->params.maxDisparity - h->params.minDisparity) / kScanLanes;
        if (!valid_schedule(h->schedule, K, why)) return h->fail(LUMEN_EINVAL, "schedule: " + why);
    }
    return LUMEN_OK;
}

int prepare(lumen_sgbm* h, int W, int H) {
    int rc = h->ensure_ready();
    if (rc) return rc;
    Geometry& g = h->geo;
    const lumen_stereo_params& p = h->params;
    g.params = p;
    g.W = W;
    g.H = H;
    g.m = p.minDisparity;
    g.D = p.maxDisparity - p.minDisparity;
    g.Dp = padded_range(g.D);
    g.npaths = num_paths(p);
    g.costScale = cost_scale(p);
    const size_t n = g.pixels();
    struct {
        DevBuf* b;
        size_t bytes;
        const char* name;
    } plan[] = {
        {&h->left, n, "left image"},
        {&h->right, n, "right image"},
        {&h->cost, g.volume(), "cost volume"},
        {&h->paths, g.volume() * 2 * g.npaths, "path costs"},
        {&h->dispL, n * 2, "left disparity"},
        {&h->dispR, n * 2, "right disparity"},
        {&h->dispLR, n * 2, "checked disparity"},
        {&h->wtaInt, n * 2, "integer disparity"},
        {&h->spk, n * 2, "speckle output"},
        {&h->out, n * 2, "output"},
        {&h->trash, kTrashElems * 2, "trash"},
    };
    for (auto& e : plan) {
        if (e.bytes == 0) continue;
        hipError_t err = e.b->ensure(e.bytes);
        if (err != hipSuccess) return h->hip_fail(err, e.name);
    }
    return LUMEN_OK;
}

// The PairJob of the last prepared frame (its parameters and size) over the handle's buffers.
PairJob make_job(const lumen_sgbm* h, const uint8_t* dl, const uint8_t* dr, int16_t* dout) {
    const Geometry& g = h->geo;
    PairJob j{};
    fill_job(j, g.params, g.W, g.H);
    if (h->hasSchedule) j.sched = h->schedule;  // checked by validate()
    j.left = dl;
    j.right = dr;
    j.cost = h->cost.as<uint8_t>();
    j.L = h->paths.as<uint16_t>();
    j.dispL = h->dispL.as<int16_t>();
    j.dispR = h->dispR.as<int16_t>();
    j.dispLR = h->dispLR.as<int16_t>();
    j.wtaInt = h->wtaInt.as<int16_t>();
    j.spk = h->spk.as<int16_t>();
    j.out = dout;
    return j;
}

// Enqueues every stage after the upload; images are already in h->left / h->right or in the
// caller's device buffers. The stage times come from events the stage kernels carry
// (StageEvents), so they are the kernels' own durations and nothing is enqueued between them.
int run_stages(lumen_sgbm* h, const uint8_t* dl, const uint8_t* dr, int16_t* dout, int slot) {
    hipStream_t s = h->stream;
    const PairJob j = make_job(h, dl, dr, dout);
    // Each event costs host time and stream time: only those of the handle's timing level.
    hipEvent_t ev[kNumEvents] = {};
    h->slotTiming[slot] = h->timing;
    if (h->timing >= LUMEN_TIMING_TOTAL) {
        ev[1] = h->ev[slot][1];
        ev[8] = h->ev[slot][8];
    }
    if (h->timing >= LUMEN_TIMING_STAGES)
        for (int k = 2; k < 8; ++k) ev[k] = h->ev[slot][k];
    hipError_t e = launch_cost(j, s, {ev[1], ev[2]});
    if (e != hipSuccess) return h->hip_fail(e, "cost launch");
    e = launch_paths(j, h->trash.as<uint16_t>(), s, {ev[3], ev[4]});
    if (e != hipSuccess) return h->hip_fail(e, "path aggregation launch");
    e = launch_wta(j, s, nullptr, {ev[5], ev[6]});
    if (e != hipSuccess) return h->hip_fail(e, "WTA launch");
    // The dump hook reads the output from the handle's own buffer: when the caller's buffer is
    // another one, the median writes both.
    e = launch_post(j, s, dout != h->out.as<int16_t>() ? h->out.as<int16_t>() : nullptr,
                    {ev[7], ev[8]});
    if (e != hipSuccess) return h->hip_fail(e, "post-processing launch");
    return LUMEN_OK;
}

int begin_slot(lumen_sgbm* h) {
    const int slot = h->ringNext;
    h->ringNext = (h->ringNext + 1) % kRing;
    h->ringCount = std::min(h->ringCount + 1, kRing);
    return slot;
}

int check_params(const lumen_stereo_params* p) {
    if (!p) {
        g_last_error = "params is NULL";
        return LUMEN_EINVAL;
    }
    return LUMEN_OK;
}

}  // namespace

void lumen::set_last_error(const std::string& msg) { g_last_error = msg; }

// ---------------------------------------------------------------------------------------
// Scratch of lumen_sgbm_compute_device_noalloc: a fixed device array that holds one pair's
// working set (cost volume, per-path costs, the intermediate maps, the aggregation's trash
// rows), so the entry point never allocates. Each device the library runs on holds its own
// copy for the life of the process. In host code `scratch` is only a handle for the symbol:
// its device address comes from hipGetSymbolAddress, once per device.
// ---------------------------------------------------------------------------------------
namespace {

constexpr size_t kNoAllocScratchBytes = size_t(192) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch (looked up on the first call per device).
hipError_t scratch_ptr(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

// Points j's intermediate buffers into the scratch at `base` (256-byte aligned pieces, as
// layout_pair does) and returns the bytes used; the trash rows are returned in `trash`.
size_t layout_scratch(PairJob& j, char* base, uint16_t** trash) {
    const size_t n = (size_t)j.W * j.H;
    size_t used = 0;
    auto take = [&](size_t bytes) {
        char* q = base ? base + used : nullptr;
        used += align_up(bytes, 256);
        return q;
    };
    j.cost = reinterpret_cast<uint8_t*>(take(j.pathStride));
    j.L = reinterpret_cast<uint16_t*>(take(j.pathStride * 2 * j.npaths));
    j.dispL = reinterpret_cast<int16_t*>(take(n * 2));
    j.dispR = reinterpret_cast<int16_t*>(take(n * 2));
    j.dispLR = reinterpret_cast<int16_t*>(take(n * 2));
    j.wtaInt = reinterpret_cast<int16_t*>(take(n * 2));
    j.spk = reinterpret_cast<int16_t*>(take(n * 2));
    *trash = reinterpret_cast<uint16_t*>(take(kTrashElems * 2));
    return used;
}

}  // namespace

extern "C" {

int lumen_sgbm_default_params(lumen_stereo_params* out) {
    if (!out) return LUMEN_EINVAL;
    out->minDisparity = 0;
    out->maxDisparity = 128;
    out->blockSize = 5;
    out->P1 = 200;
    out->P2 = 800;
    out->preFilterCap = 31;
    out->uniquenessRatio = 10;
    out->disp12MaxDiff = 1;
    out->speckleWindowSize = 100;
    out->speckleRange = 32;
    out->useGPU = 1;
    out->useSharedMemory = 1;
    out->numDirections = 4;
    out->matchingCostMode = 0;
    return LUMEN_OK;
}

int lumen_sgbm_create(const lumen_stereo_params* params, lumen_sgbm** out) {
    if (!out) {
        g_last_error = "out is NULL";
        return LUMEN_EINVAL;
    }
    *out = nullptr;
    lumen_stereo_params p;
    if (params)
        p = *params;
    else
        lumen_sgbm_default_params(&p);
    lumen_sgbm* h = new (std::nothrow) lumen_sgbm();
    if (!h) {
        g_last_error = "out of host memory";
        return LUMEN_ENOMEM;
    }
    h->params = p;
    *out = h;
    return LUMEN_OK;
}

int lumen_sgbm_destroy(lumen_sgbm* h) {
    delete h;
    return LUMEN_OK;
}

int lumen_sgbm_set_params(lumen_sgbm* h, const lumen_stereo_params* params) {
    if (!h) return LUMEN_EINVAL;
    if (check_params(params)) return h->fail(LUMEN_EINVAL, "params is NULL");
    h->params = *params;
    return LUMEN_OK;
}

int lumen_sgbm_get_params(const lumen_sgbm* h, lumen_stereo_params* out) {
    if (!h || !out) return LUMEN_EINVAL;
    *out = h->params;
    return LUMEN_OK;
}

int lumen_sgbm_num_paths(const lumen_sgbm* h) { return h ? num_paths(h->params) : 0; }

int lumen_sgbm_set_schedule(lumen_sgbm* h, const lumen_schedule* s) {
    if (!h) return LUMEN_EINVAL;
    if (!s) {
        h->hasSchedule = false;
        return LUMEN_OK;
    }
    h->schedule =
        Schedule{s->agg_prefetch, s->agg_priority, s->census_tile, s->agg_wide, s->agg_occupancy};
    h->hasSchedule = true;
    return LUMEN_OK;
}

int lumen_sgbm_get_schedule(const lumen_sgbm* h, int width, int height, lumen_schedule* out) {
    if (!h || !out || width <= 0 || height <= 0) return LUMEN_EINVAL;
    const long long D = static_cast<long long>(h->params.maxDisparity) - h->params.minDisparity;
    if (D < 1 || D > LUMEN_MAX_DISPARITY_RANGE) return LUMEN_EINVAL;
    const Schedule s = h->hasSchedule ? h->schedule
                                      : tuned_schedule(width, padded_range(static_cast<int>(D)),
                                                       h->params.matchingCostMode,
                                                       num_paths(h->params));
    out->agg_prefetch = s.aggPrefetch;
    out->agg_priority = s.aggPriority;
    out->census_tile = s.censusTile;
    out->agg_wide = s.aggWide;
    out->agg_occupancy = s.aggOccupancy;
    return LUMEN_OK;
}

int lumen_sgbm_footprint(const lumen_stereo_params* params, int width, int height,
                         size_t* bytes_out) {
    if (!params || !bytes_out || width <= 0 || height <= 0) return LUMEN_EINVAL;
    const long long D = static_cast<long long>(params->maxDisparity) - params->minDisparity;
    if (D < 1 || D > LUMEN_MAX_DISPARITY_RANGE) return LUMEN_EINVAL;
    *bytes_out = pair_footprint(*params, width, height);
    return LUMEN_OK;
}

int lumen_sgbm_compute(lumen_sgbm* h, const uint8_t* left, size_t left_stride,
                       const uint8_t* right, size_t right_stride, int width, int height,
                       int channels, int16_t* out, size_t out_stride) {
    if (!h) return LUMEN_EINVAL;
    h->err.clear();
    int rc = validate(h, left, right, width, height);
    if (rc) return rc;
    if (!out) return h->fail(LUMEN_EINVAL, "Output buffer is NULL");
    rc = prepare(h, width, height);
    if (rc) return rc;
    h->haveResult = false;
    const int W = width, H = height;
    const int ch = channels == 3 ? 3 : 1;
    if (left_stride == 0) left_stride = (size_t)W * ch;
    if (right_stride == 0) right_stride = (size_t)W * ch;
    if (out_stride == 0) out_stride = (size_t)W * 2;
    const size_t n = (size_t)W * H;
    hipStream_t s = h->stream;
    // The previous call's download may still be reading the staging buffers only if it failed
    // half-way; wait for the stream before the host writes them.
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return h->hip_fail(e, "stream");
    e = h->stageIn.ensure(2 * n);
    if (e != hipSuccess) return h->hip_fail(e, "pinned input staging");
    e = h->stageOut.ensure(2 * n);
    if (e != hipSuccess) return h->hip_fail(e, "pinned output staging");
    uint8_t* pin = h->stageIn.as<uint8_t>();
    if (ch == 3) {
        // The conversion src/sgbm_gpu.cu:168-175 applies to 3-channel input (OpenCV's BGR2GRAY).
        parallel_bgr_to_gray(pin, left, left_stride, W, H);
        parallel_bgr_to_gray(pin + n, right, right_stride, W, H);
    } else if (left_stride == (size_t)W && right_stride == (size_t)W) {
        // Packed images: one parallel copy of both, so a 1080p pair (4 MB) uses the pool.
        parallel_copy({{pin, left, n}, {pin + n, right, n}});
    } else {
        parallel_copy_rows(pin, W, left, left_stride, W, H);
        parallel_copy_rows(pin + n, W, right, right_stride, W, H);
    }
    const int slot = begin_slot(h);
    if (h->timing >= LUMEN_TIMING_TOTAL) {
        e = hipEventRecord(h->ev[slot][0], s);
        if (e != hipSuccess) return h->hip_fail(e, "event record");
    }
    e = hipMemcpyAsync(h->left.ptr, pin, n, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return h->hip_fail(e, "upload left");
    e = hipMemcpyAsync(h->right.ptr, pin + n, n, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return h->hip_fail(e, "upload right");
    h->slotHasUpload[slot] = true;
    rc = run_stages(h, h->left.as<uint8_t>(), h->right.as<uint8_t>(), h->out.as<int16_t>(), slot);
    if (rc) return rc;
    e = hipMemcpyAsync(h->stageOut.ptr, h->out.ptr, n * 2, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return h->hip_fail(e, "download");
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return h->hip_fail(e, "compute");
    parallel_copy_rows(out, out_stride, h->stageOut.ptr, (size_t)W * 2, (size_t)W * 2, H);
    h->haveResult = true;
    return LUMEN_OK;
}

int lumen_sgbm_compute_device(lumen_sgbm* h, const uint8_t* d_left, const uint8_t* d_right,
                              int width, int height, int16_t* d_out, int sync) {
    if (!h) return LUMEN_EINVAL;
    h->err.clear();
    int rc = validate(h, d_left, d_right, width, height);
    if (rc) return rc;
    if (!d_out) return h->fail(LUMEN_EINVAL, "Output buffer is NULL");
    rc = prepare(h, width, height);
    if (rc) return rc;
    h->haveResult = false;
    const int slot = begin_slot(h);
    hipStream_t s = h->stream;
    h->slotHasUpload[slot] = false;  // the compute starts with the cost kernel (event 1)
    rc = run_stages(h, d_left, d_right, d_out, slot);
    if (rc) return rc;
    if (sync) {
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) return h->hip_fail(e, "compute");
    }
    h->haveResult = true;
    return LUMEN_OK;
}

int lumen_sgbm_compute_device_noalloc(const lumen_stereo_params* params, const uint8_t* d_left,
                                      const uint8_t* d_right, int width, int height,
                                      int16_t* d_out, void* stream) {
    if (check_params(params)) return LUMEN_EINVAL;
    const std::string msg = validate_pair(*params, width, height, d_left && d_right);
    if (!msg.empty()) {
        g_last_error = msg;
        return LUMEN_EINVAL;
    }
    if (!d_out) {
        g_last_error = "Output buffer is NULL";
        return LUMEN_EINVAL;
    }
    PairJob j{};
    fill_job(j, *params, width, height);
    uint16_t* trash = nullptr;
    const size_t need = layout_scratch(j, nullptr, &trash);
    if (need > kNoAllocScratchBytes) {
        g_last_error = "pair needs " + std::to_string(need) +
                       " bytes of working memory; lumen_sgbm_compute_device_noalloc has " +
                       std::to_string(kNoAllocScratchBytes) +
                       " (W*H*Dp*(1 + 2*paths) plus five int16 maps)";
        return LUMEN_EINVAL;
    }
    void* base = nullptr;
    hipError_t e = scratch_ptr(&base);
    if (e != hipSuccess) {
        g_last_error = std::string("scratch address: ") + hipGetErrorString(e);
        return LUMEN_EDEVICE;
    }
    layout_scratch(j, static_cast<char*>(base), &trash);
    j.left = d_left;
    j.right = d_right;
    j.out = d_out;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const char* what = "cost launch";
    e = launch_cost(j, s);
    if (e == hipSuccess) {
        what = "path aggregation launch";
        e = launch_paths(j, trash, s);
    }
    if (e == hipSuccess) {
        what = "WTA launch";
        e = launch_wta(j, s);
    }
    if (e == hipSuccess) {
        what = "post-processing launch";
        e = launch_post(j, s);
    }
    if (e != hipSuccess) {
        g_last_error = std::string(what) + ": " + hipGetErrorString(e);
        return LUMEN_EDEVICE;
    }
    return LUMEN_OK;
}

int lumen_sgbm_set_timing(lumen_sgbm* h, int level) {
    if (!h) return LUMEN_EINVAL;
    if (level < LUMEN_TIMING_NONE || level > LUMEN_TIMING_STAGES)
        return h->fail(LUMEN_EINVAL, "timing level must be 0, 1 or 2");
    h->timing = level;
    return LUMEN_OK;
}

int lumen_sgbm_get_stream(lumen_sgbm* h, void** stream_out) {
    if (!h || !stream_out) return LUMEN_EINVAL;
    int rc = h->ensure_ready();
    if (rc) return rc;
    *stream_out = static_cast<void*>(h->stream);
    return LUMEN_OK;
}

static int read_slot(lumen_sgbm* h, int slot, float out[6]) {
    for (int i = 0; i < 6; ++i) out[i] = 0.0f;
    const int level = h->slotTiming[slot];
    if (level == LUMEN_TIMING_NONE) {
        hipError_t e = hipStreamSynchronize(h->stream);
        return e == hipSuccess ? LUMEN_OK : h->hip_fail(e, "compute");
    }
    hipError_t e = hipEventSynchronize(h->ev[slot][8]);
    if (e != hipSuccess) return h->hip_fail(e, "event sync");
    // Without an upload, the compute starts at event 1 and the upload time is 0.
    const int s0 = h->slotHasUpload[slot] ? 0 : 1;
    const int pairs[6][2] = {{s0, 8}, {1, 2}, {3, 4}, {7, 8}, {5, 6}, {s0, 1}};
    for (int i = 0; i < 6; ++i) {
        float ms = 0.0f;
        const bool recorded = level >= LUMEN_TIMING_STAGES || i == 0 || i == 5;
        if (recorded && pairs[i][0] != pairs[i][1]) {
            e = hipEventElapsedTime(&ms, h->ev[slot][pairs[i][0]], h->ev[slot][pairs[i][1]]);
            if (e != hipSuccess) return h->hip_fail(e, "event elapsed");
        }
        out[i] = ms;
    }
    return LUMEN_OK;
}

int lumen_sgbm_get_timings(lumen_sgbm* h, float out[6]) {
    if (!h || !out) return LUMEN_EINVAL;
    if (!h->ready || h->ringCount == 0) {
        for (int i = 0; i < 6; ++i) out[i] = 0.0f;
        return LUMEN_OK;
    }
    const int slot = (h->ringNext + kRing - 1) % kRing;
    return read_slot(h, slot, out);
}

int lumen_sgbm_get_timings_avg(lumen_sgbm* h, int ncalls, float out[6]) {
    if (!h || !out || ncalls < 1) return LUMEN_EINVAL;
    for (int i = 0; i < 6; ++i) out[i] = 0.0f;
    if (!h->ready || h->ringCount == 0) return LUMEN_OK;
    const int n = std::min(ncalls, h->ringCount);
    for (int k = 1; k <= n; ++k) {
        float t[6];
        int rc = read_slot(h, (h->ringNext + kRing - k) % kRing, t);
        if (rc) return rc;
        for (int i = 0; i < 6; ++i) out[i] += t[i] / n;
    }
    return LUMEN_OK;
}

const char* lumen_sgbm_last_error(const lumen_sgbm* h) { return h ? h->err.c_str() : ""; }

const char* lumen_last_error(void) { return g_last_error.c_str(); }

int lumen_sgbm_dump_stage(lumen_sgbm* h, int stage, int path, void* host, size_t bytes) {
    if (!h || !host) return LUMEN_EINVAL;
    if (!h->haveResult) return h->fail(LUMEN_EINVAL, "no completed compute to dump");
    const Geometry& g = h->geo;
    const size_t n = g.pixels();
    hipStream_t s = h->stream;
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return h->hip_fail(e, "sync");
    auto strip = [&](const void* dev, size_t elem) -> int {
        // Padded [pixels][Dp] -> host [pixels][D]: stripped on the device, then one copy.
        const size_t need = n * g.D * elem;
        if (bytes < need) return h->fail(LUMEN_EINVAL, "dump buffer too small");
        const void* src = dev;
        if (g.D != g.Dp) {
            hipError_t err = h->dumpTmp.ensure(need);
            if (err != hipSuccess) return h->hip_fail(err, "dump buffer");
            err = launch_strip(dev, h->dumpTmp.ptr, n, g.D, g.Dp, static_cast<int>(elem), s);
            if (err != hipSuccess) return h->hip_fail(err, "dump strip");
            err = hipStreamSynchronize(s);
            if (err != hipSuccess) return h->hip_fail(err, "dump strip");
            src = h->dumpTmp.ptr;
        }
        hipError_t err = hipMemcpy(host, src, need, hipMemcpyDeviceToHost);
        if (err != hipSuccess) return h->hip_fail(err, "dump copy");
        return LUMEN_OK;
    };
    auto plain = [&](const void* dev) -> int {
        if (bytes < n * 2) return h->fail(LUMEN_EINVAL, "dump buffer too small");
        hipError_t err = hipMemcpy(host, dev, n * 2, hipMemcpyDeviceToHost);
        if (err != hipSuccess) return h->hip_fail(err, "dump copy");
        return LUMEN_OK;
    };
    switch (stage) {
        case LUMEN_STAGE_COST: return strip(h->cost.ptr, 1);
        case LUMEN_STAGE_PATH:
            if (path < 0 || path >= g.npaths) return h->fail(LUMEN_EINVAL, "path index out of range");
            return strip(h->paths.as<uint16_t>() + (size_t)path * g.volume(), 2);
        case LUMEN_STAGE_SUM: {
            // The sum the WTA works on lives only in registers: the WTA rows run again (same
            // inputs, same maps written) in the instantiation that also stores it.
            e = h->sum.ensure(g.volume() * 2);
            if (e != hipSuccess) return h->hip_fail(e, "sum buffer");
            const PairJob j = make_job(h, nullptr, nullptr, h->out.as<int16_t>());
            e = launch_wta(j, s, h->sum.as<uint16_t>());
            if (e != hipSuccess) return h->hip_fail(e, "sum launch");
            e = hipStreamSynchronize(s);
            if (e != hipSuccess) return h->hip_fail(e, "sum");
            return strip(h->sum.ptr, 2);
        }
        case LUMEN_STAGE_WTA_INT: return plain(h->wtaInt.ptr);
        case LUMEN_STAGE_DISP_LEFT: return plain(h->dispL.ptr);
        case LUMEN_STAGE_DISP_RIGHT: return plain(h->dispR.ptr);
        case LUMEN_STAGE_AFTER_LR: return plain(h->dispLR.ptr);
        case LUMEN_STAGE_AFTER_SPECKLE:
            return plain(g.params.speckleWindowSize > 0 ? h->spk.ptr : h->dispLR.ptr);
        case LUMEN_STAGE_OUTPUT: return plain(h->out.ptr);
        default: return h->fail(LUMEN_EINVAL, "unknown stage");
    }
}

}  // extern "C"
