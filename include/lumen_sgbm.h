/*
 * lumen_sgbm.h - C ABI of the MI355X SGBM matcher (liblumen_sgbm.so).
 *
 * This is the drop-in boundary for the reference's matcher. Everything is extern "C", takes
 * plain pointers and sizes, never throws, and returns an int status:
 *   LUMEN_OK (0), LUMEN_EINVAL (1), LUMEN_EDEVICE (2), LUMEN_ENOMEM (3), and
 *   LUMEN_ETIMEDOUT (4) from lumen_server_wait only.
 * The text of the last error of a handle is returned by lumen_sgbm_last_error(); errors that
 * occur before a handle exists are reported by lumen_last_error().
 *
 * Threading: a lumen_sgbm handle, like the reference's StereoSGBM (one stream, member buffers),
 * must not be used by two threads at once; separate handles are independent. The same holds
 * for a lumen_batch handle. A lumen_server may be submitted to and waited on from any number of
 * threads (it locks internally). The tuning table is process-wide and locked.
 *
 * Citations are into the reference (Pfactorial01/LumenStereo); each entry point names the
 * interface it replaces.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LUMEN_OK 0
#define LUMEN_EINVAL 1
#define LUMEN_EDEVICE 2
#define LUMEN_ENOMEM 3
#define LUMEN_ETIMEDOUT 4

/* Largest disparity range (maxDisparity - minDisparity); replaces kSgbmMaxDisparityRange
 * (include/stereo/sgbm_limits.h:9). */
#define LUMEN_MAX_DISPARITY_RANGE 512

/* Mirrors stereo::StereoParams field by field (include/stereo/stereo_params.h:57-82) as int32.
 * useGPU is accepted and ignored, as in the reference (nothing reads it). useSharedMemory is
 * accepted and does not change results (src/sgbm_gpu.cu:235-285 gives identical costs either
 * way). matchingCostMode: 0 = SAD, 1 = Census (include/stereo/stereo_params.h:20-23). */
typedef struct lumen_stereo_params {
    int32_t minDisparity;
    int32_t maxDisparity;
    int32_t blockSize;
    int32_t P1;
    int32_t P2;
    int32_t preFilterCap;
    int32_t uniquenessRatio;
    int32_t disp12MaxDiff;
    int32_t speckleWindowSize;
    int32_t speckleRange;
    int32_t useGPU;
    int32_t useSharedMemory;
    int32_t numDirections;
    int32_t matchingCostMode;
} lumen_stereo_params;

typedef struct lumen_sgbm lumen_sgbm;

/* Writes the StereoParams defaults (include/stereo/stereo_params.h:59-82). */
int lumen_sgbm_default_params(lumen_stereo_params* out);

/* Replaces StereoSGBM::StereoSGBM(const StereoParams&) (src/sgbm_gpu.cu:100-107). No device
 * call is made until the first compute; the stream and events are created then. */
int lumen_sgbm_create(const lumen_stereo_params* params, lumen_sgbm** out);

/* Replaces StereoSGBM::~StereoSGBM (src/sgbm_gpu.cu:109-113). NULL is accepted. */
int lumen_sgbm_destroy(lumen_sgbm* h);

/* Replace StereoSGBM::setParams / getParams (src/sgbm_gpu.cu:115-117,
 * include/stereo/sgbm_gpu.h:47). */
int lumen_sgbm_set_params(lumen_sgbm* h, const lumen_stereo_params* params);
int lumen_sgbm_get_params(const lumen_sgbm* h, lumen_stereo_params* out);

/* Replaces StereoSGBM::compute(const cv::Mat&, const cv::Mat&, cv::Mat&)
 * (src/sgbm_gpu.cu:147-195). Host images of width x height with `channels` 1 (grayscale) or 3
 * (BGR, converted to gray on the host as the reference's cvtColor does); any other channel
 * count is treated as grayscale bytes. Row strides are in bytes (0 = packed). Writes int16
 * disparity in 1/16 px (-1 = invalid) to `out` (stride in bytes, 0 = packed). Blocking.
 * Validation order and messages follow src/sgbm_gpu.cu:148-166. */
int lumen_sgbm_compute(lumen_sgbm* h, const uint8_t* left, size_t left_stride,
                       const uint8_t* right, size_t right_stride, int width, int height,
                       int channels, int16_t* out, size_t out_stride);

/* Device-resident form of lumen_sgbm_compute (no reference counterpart: the reference always
 * copies from host memory). d_left/d_right are packed grayscale images already in device
 * memory, d_out a packed int16 device buffer. Work is enqueued on the handle's stream; when
 * `sync` is non-zero the call waits for it. */
int lumen_sgbm_compute_device(lumen_sgbm* h, const uint8_t* d_left, const uint8_t* d_right,
                              int width, int height, int16_t* d_out, int sync);

/* Allocation-free form of lumen_sgbm_compute_device. It has no counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). It takes the parameters instead of a handle
 * and enqueues the same stages on `stream` (a hipStream_t; NULL = the null stream) without
 * waiting; it calls no hipMalloc, hipFree or other allocating routine. The pair's working set
 * (W*H*Dp*(1 + 2*paths) bytes, Dp = D rounded up to 64, plus five int16 maps) lives in a fixed
 * 192 MiB device array of the library, one per GPU, held until the process ends: 640x480 D=64
 * with 4 paths fits, 1920x1080 D=128 does not; a larger pair fails with LUMEN_EINVAL and
 * lumen_last_error() gives the sizes. Every call on one GPU uses that GPU's array, so calls
 * that overlap on one GPU (other threads or streams before this call's work has finished)
 * race and give undefined results. No stage reads scratch it has not written in the same
 * call, so nothing is cleared between calls. Validation and messages are those of
 * lumen_sgbm_compute; the schedule is the tuned one. */
int lumen_sgbm_compute_device_noalloc(const lumen_stereo_params* params, const uint8_t* d_left,
                                      const uint8_t* d_right, int width, int height,
                                      int16_t* d_out, void* stream);

/* The hipStream_t the handle launches on (as void*), creating it if needed. */
int lumen_sgbm_get_stream(lumen_sgbm* h, void** stream_out);

/* Stage times of the last completed compute in milliseconds, from hipEvents on the handle's
 * stream: out[0] total (as getLastComputeTimeMs, src/sgbm_gpu.cu:177-191: upload through
 * post-processing; for lumen_sgbm_compute_device from the first kernel's start), out[1] cost
 * (getCostComputeTimeMs), out[2] aggregation (getAggregationTimeMs), out[3] post-processing
 * (getPostProcessTimeMs: speckle and median), out[4] path sum + WTA + LR check, out[5] upload.
 * The events are carried by the stage kernels' dispatches (hipExtLaunchKernel), so the stage
 * times are the kernels' own durations; recording them still costs time (lumen_sgbm_set_timing
 * turns them off). Waits for the last compute. */
int lumen_sgbm_get_timings(lumen_sgbm* h, float out[6]);

/* Average of the same six stage times over the last `ncalls` completed computes (at most 64
 * are kept). Waits for them. */
int lumen_sgbm_get_timings_avg(lumen_sgbm* h, int ncalls, float out[6]);

/* What later computes on the handle time (no reference counterpart: the reference always times
 * every stage). LUMEN_TIMING_STAGES (the default) records every stage time above;
 * LUMEN_TIMING_TOTAL only out[0] and out[5] (getLastComputeTimeMs), the others read 0;
 * LUMEN_TIMING_NONE nothing (all read 0). Each stage event costs host time per call and stream
 * time between the kernels (about 20 us of a 640x480 D=64 frame), so callers that do not read
 * the stage times can turn them off. */
#define LUMEN_TIMING_NONE 0
#define LUMEN_TIMING_TOTAL 1
#define LUMEN_TIMING_STAGES 2
int lumen_sgbm_set_timing(lumen_sgbm* h, int level);

/* Text of the last error on this handle ("" if none). Valid until the next call on it. */
const char* lumen_sgbm_last_error(const lumen_sgbm* h);

/* Text of the last error raised without a handle (e.g. by lumen_sgbm_create), per thread. */
const char* lumen_last_error(void);

/* Parity hook (no reference counterpart): copies one stage of the last compute into host
 * memory in the reference's layout (idx = (y*W + x)*D + d, D = maxDisparity - minDisparity).
 * Stages: */
#define LUMEN_STAGE_COST 0        /* uint8  W*H*D */
#define LUMEN_STAGE_PATH 1        /* uint16 W*H*D, path index in `path` (0..npaths-1) */
#define LUMEN_STAGE_SUM 2         /* uint16 W*H*D, sum of all paths mod 2^16 */
#define LUMEN_STAGE_WTA_INT 3     /* int16  W*H, integer best disparity index of the left WTA */
#define LUMEN_STAGE_DISP_LEFT 4   /* int16  W*H, left WTA in 1/16 px before the LR check */
#define LUMEN_STAGE_DISP_RIGHT 5  /* int16  W*H, right WTA in 1/16 px */
#define LUMEN_STAGE_AFTER_LR 6    /* int16  W*H, after the LR check */
#define LUMEN_STAGE_AFTER_SPECKLE 7 /* int16 W*H, after the speckle filter */
#define LUMEN_STAGE_OUTPUT 8      /* int16  W*H, final output (after the median) */
/* Path order: 0 LR, 1 RL, 2 TB, 3 BT, 4 TLBR, 5 BRTL, 6 TRBL, 7 BLTR
 * (the reference's launch order, src/sgbm_gpu.cu:321-421). */
int lumen_sgbm_dump_stage(lumen_sgbm* h, int stage, int path, void* host, size_t bytes);

/* Number of SGM paths the handle's parameters select (4, or 8 when numDirections >= 8;
 * src/sgbm_gpu.cu:359). */
int lumen_sgbm_num_paths(const lumen_sgbm* h);

/* Schedule of the per-pair path (no reference counterpart): how the work of a pair is grouped
 * and ordered on the device, never what is computed; every valid schedule gives bit-identical
 * results. By default each pair runs the schedule the tuning table names for its width, padded
 * disparity range (D rounded up to 64), cost mode and path count (see lumen_tuning_load). */
typedef struct lumen_schedule {
    int32_t agg_prefetch;  /* aggregation cost prefetch depth: 4, 8 or 16 (16 only where a
                              block's stores fit: padded range <= 192 or 256, 384) */
    int32_t agg_priority;  /* aggregation wave priority: 0 none, 1 by path kind, 2 by length */
    int32_t census_tile;   /* census cost segment width in pixels: 128, 256 or 512 */
    int32_t agg_wide;      /* aggregation paths whose scanlines span 32 lanes (two per wave,
                              half the disparities per lane: a shorter dependency chain per
                              step for twice the waves) instead of 16 (four per wave): 0 none,
                              1 the straight pair along the longer image dimension, 2 all four
                              straight paths, 3 every path; 4, 5 and 6 the same sets on 64-lane
                              scanlines (one per wave, a quarter of the disparities per lane;
                              padded ranges that are a multiple of 128 only). */
    int32_t agg_occupancy; /* aggregation blocks (of four waves) per CU at most: 0 no cap, or
                              2, 3 or 4 (the kernel then reserves LDS it does not use). Fewer
                              concurrent scanlines per CU can move a large frame's bytes
                              faster. */
} lumen_schedule;

/* Overrides the tuned schedule for this handle (NULL restores it). Checked at compute time:
 * a schedule the pair's range does not support fails with LUMEN_EINVAL. */
int lumen_sgbm_set_schedule(lumen_sgbm* h, const lumen_schedule* s);

/* The schedule a width x height pair would run with on this handle. */
int lumen_sgbm_get_schedule(const lumen_sgbm* h, int width, int height, lumen_schedule* out);

/* Replaces the process-wide tuning table with the file at `path` (NULL or "" = tuning.tsv next
 * to the library, which is also read on first use). On error the table is unchanged and
 * lumen_last_error() says why. entries_out (may be NULL) receives the number of entries. */
int lumen_tuning_load(const char* path, int* entries_out);

/* Empties the tuning table: every pair runs the default schedule. */
int lumen_tuning_clear(void);

/* Device bytes the handle needs for a width x height pair with its current parameters:
 * the padded cost volume, the per-path uint16 buffers and the image-sized buffers. */
int lumen_sgbm_footprint(const lumen_stereo_params* params, int width, int height,
                         size_t* bytes_out);

/* Deterministic synthetic stereo pair (no reference counterpart; the benchmark scenes):
 * band-limited noise left image, slanted planes plus fronto-parallel occluders as ground
 * truth with disparities in [dmin, dmax], right image = left warped by -d with occlusions
 * filled by fresh noise, gain/bias/noise applied. lowtex_frac of the area gets its texture
 * std reduced to 5. gt may be NULL. The same seed gives the same bytes on every machine. */
int lumen_synth_pair(int width, int height, float dmin, float dmax, uint64_t seed,
                     float lowtex_frac, uint8_t* left, uint8_t* right, float* gt);

/* ------------------------------------------------------------------------------------------
 * Batched form (no reference counterpart). Many pairs of different size, disparity range, cost
 * mode and path count are processed stage by stage in persistent kernels over work lists, the
 * whole pipeline replayed as one hipGraph. Every pair's disparity is bit-identical to
 * lumen_sgbm_compute of the same pair with the same parameters.
 * ---------------------------------------------------------------------------------------- */
typedef struct lumen_pair_desc {
    lumen_stereo_params params;
    int32_t width;
    int32_t height;
    const uint8_t* left;   /* packed grayscale, host or device memory (see device_io) */
    const uint8_t* right;
    int16_t* out;          /* packed int16 disparity, host or device memory */
} lumen_pair_desc;

typedef struct lumen_batch lumen_batch;

/* A batch executor with a device arena of arena_bytes (0 = half of the free device memory) for
 * the pairs' working sets; a batch larger than the arena runs as several sub-batches. */
int lumen_batch_create(size_t arena_bytes, lumen_batch** out);
int lumen_batch_destroy(lumen_batch* b);
int lumen_batch_arena_bytes(const lumen_batch* b, size_t* out);

/* Runs npairs pairs. device_io = 0: images and outputs are host memory (staged through pinned
 * buffers); device_io = 1: they are device memory. Blocking. Pairs are validated as
 * lumen_sgbm_compute validates them. */
int lumen_batch_run(lumen_batch* b, const lumen_pair_desc* pairs, int npairs, int device_io);

/* Times of the last run in ms: out[0] total (host wall clock of lumen_batch_run), out[1]
 * upload, out[2] compute, out[3] download. A host-memory batch of 4 or more pairs runs as 2 or 4
 * chunks alternating over two streams, so that one chunk's staging and transfers overlap another
 * chunk's kernels; out[1..3] are then sums over the chunks and may add up to more than out[0]. */
int lumen_batch_get_timings(const lumen_batch* b, float out[4]);

const char* lumen_batch_last_error(const lumen_batch* b);

/* ------------------------------------------------------------------------------------------
 * Serving form (no reference counterpart). Requests (stereo pairs of any size, range, cost mode
 * and path count) are submitted as they arrive and return a ticket. Requests are admitted into
 * a device pool by their working-set size (cost volume, path buffers, maps), and whenever one
 * of the lanes (a stream with its own persistent-kernel pipeline) is free, admitted requests of
 * similar size start on it as one batch: from the submitting thread when a lane is idle and
 * nothing is waiting, otherwise from a dispatcher thread. So requests that arrive while others
 * are on the GPU run alongside them. Host images are staged into a pinned block of the server
 * by the submitting thread, and a waiter copies its output from the pinned block itself, so
 * the dispatcher never copies host memory. Every result is bit-identical to
 * lumen_sgbm_compute of the same pair with the same parameters.
 * ---------------------------------------------------------------------------------------- */
typedef struct lumen_server_config {
    size_t pool_bytes;   /* device pool; 0 = half of the free device memory */
    int32_t lanes;       /* batches in flight at once, 1..4; 0 = 2 */
    int32_t max_batch;   /* most requests one batch takes; 0 = 64 */
    int32_t device_io;   /* 0: request images/outputs are host memory; 1: device memory */
    int32_t max_bypass;  /* smaller requests that may pass a waiting one; 0 = 8, -1 = none */
    int32_t batch_budget_us; /* work one batch may hold, in microseconds at 5 TB/s of its
                                algorithmic HBM traffic (W*H*Dp*(1+5*paths) bytes); 0 = 2000.
                                A batch only groups requests within 4x of each other's work, so
                                small requests do not wait for a large one. */
    int32_t pinned_mb;   /* pinned host staging pool in MiB (host-memory requests); 0 = 1024.
                            A request that finds no room there is staged by the dispatcher
                            through its lane's own buffers instead. */
    int32_t large_us;    /* requests whose work (as for batch_budget_us) is above this many
                            microseconds are large: they run one at a time, in arrival order, on
                            lane 0, which takes other requests only while no large one is ready;
                            the other lanes never take a large one. 0 = no large class. */
    int32_t reserved;
} lumen_server_config;

typedef struct lumen_request_info {
    int32_t status;      /* LUMEN_OK or the error code of the request */
    int32_t lane;        /* lane it ran on */
    int32_t batch_size;  /* requests in its batch */
    int32_t reserved;
    int64_t t_submit_ns; /* CLOCK_MONOTONIC times (lumen_now_ns) */
    int64_t t_admit_ns;  /* pool range granted */
    int64_t t_start_ns;  /* its batch enqueued */
    int64_t t_done_ns;   /* output in the caller's buffer */
} lumen_request_info;

typedef struct lumen_server_stats {
    uint64_t submitted, completed, failed, batches;
    uint64_t max_batch_seen;
    uint64_t max_lanes_busy;   /* most lanes that had a batch in flight at the same time */
    uint64_t pool_bytes, pool_high_water;
    uint64_t max_waiting;      /* most requests waiting for admission at once */
    uint64_t direct_starts;    /* batches started by a submitting thread on an idle lane */
    uint64_t unstaged;         /* host requests that found no room in the pinned pool */
} lumen_server_stats;

typedef struct lumen_server lumen_server;

/* CLOCK_MONOTONIC in ns (the clock of lumen_request_info). */
int64_t lumen_now_ns(void);

int lumen_server_create(const lumen_server_config* cfg, lumen_server** out);
/* Finishes every submitted request, then frees the server. */
int lumen_server_destroy(lumen_server* s);
/* Validates the pair as lumen_sgbm_compute does and queues it; never blocks on the device.
 * LUMEN_ENOMEM if its working set is larger than the whole pool. The images and the output must
 * stay valid until lumen_server_wait returns for the ticket. */
int lumen_server_submit(lumen_server* s, const lumen_pair_desc* pair, uint64_t* ticket);
/* Waits for a ticket (timeout_us < 0: no limit) and releases it. Returns the request's status
 * (the ticket is then released and the request's buffers belong to the caller again), or
 * LUMEN_EINVAL for an unknown ticket or one that another thread is already waiting on. On
 * timeout it returns LUMEN_ETIMEDOUT and keeps the ticket: the request may still be queued or
 * running, so its images and output must stay valid until a later wait returns its status.
 * A ticket has at most one waiter at a time. info may be NULL. */
int lumen_server_wait(lumen_server* s, uint64_t ticket, int64_t timeout_us,
                      lumen_request_info* info);
/* Waits until no request is queued or running. */
int lumen_server_drain(lumen_server* s);
int lumen_server_get_stats(lumen_server* s, lumen_server_stats* out);
const char* lumen_server_last_error(const lumen_server* s);

#ifdef __cplusplus
}
#endif
