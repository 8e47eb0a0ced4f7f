// batch_engine.h - the persistent-kernel pipeline of the batched form, on one stream.
//
// A BatchEngine owns the device work lists, the pair table, the work counters and the
// hipGraph of the stage kernels. enqueue() takes pairs whose PairJob pointers already address
// device memory (images, cost volume, path buffers, disparity maps), writes the tables and
// launches the graph on the engine's stream; it does not wait. The caller must not call
// enqueue() again before the previous run has finished (the tables are rewritten); idle()
// tells whether it has.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "sgbm_internal.h"

namespace lumen {

// Device-resident prefix tables of a batch, each with room for `cap` pairs: costStart,
// wta0Start, wta1Start, spkStart, medStart ([cap+1] each), then wta0Pairs, wta1Pairs, spkPairs
// ([cap] each); hdr holds the batch's sizes. Unit u of a stage belongs to the pair p with
// start[p] <= u < start[p+1].
struct Tables {
    const int* hdr;
    const int* costStart;
    const int* wta0Start;
    const int* wta1Start;
    const int* spkStart;
    const int* medStart;
    const int* wta0Pairs;
    const int* wta1Pairs;
    const int* spkPairs;
};

// One aggregation unit: four consecutive scanlines of one path of one pair.
struct PathItem {
    uint16_t pair;
    uint8_t path;
    uint8_t K;
    int32_t s0;
};

// Host-side plan of one batch; fill `jobs`, the rest is derived by enqueue().
struct BatchPlan {
    std::vector<PairJob> jobs;
    std::vector<PathItem> items, itemsBig;
    std::vector<int> costStart, wta0Start, wta1Start, spkStart, medStart;
    std::vector<int> wta0Pairs, wta1Pairs, spkPairs;
    std::vector<int> special;  // pairs whose cost needs the per-pair kernels
    size_t costLds = 0, wta0Lds = 0, wta1Lds = 0;
};

// Fills everything in `pl` but `jobs` from the jobs (work lists sorted longest first).
void build_batch_plan(BatchPlan& pl);

class BatchEngine {
public:
    BatchEngine() = default;
    ~BatchEngine();
    BatchEngine(const BatchEngine&) = delete;
    BatchEngine& operator=(const BatchEngine&) = delete;

    // Creates the stream (default priority, non-blocking) and the fixed device buffers.
    hipError_t init();
    hipStream_t stream() const { return stream_; }

    // Enqueues the whole pipeline for pl.jobs (at most 65535 pairs). On failure returns a
    // LUMEN_* code with the reason in `err`.
    int enqueue(BatchPlan& pl, std::string& err);

    // True when the last enqueued run has completed (or none was enqueued).
    bool idle() const;

private:
    int ensure_graph(const BatchPlan& pl, std::string& err);

    hipStream_t stream_ = nullptr;
    hipEvent_t done_ = nullptr;
    bool pending_ = false;
    PairJob* jobs_ = nullptr;
    size_t jobsCap_ = 0;
    PathItem* items_ = nullptr;
    size_t itemsCap_ = 0;
    PathItem* itemsBig_ = nullptr;
    size_t itemsBigCap_ = 0;
    int* tables_ = nullptr;
    size_t pairCap_ = 0;
    int* ctr_ = nullptr;
    uint16_t* trash_ = nullptr;
    // Pinned mirror the tables are uploaded from (it must outlive the asynchronous copies).
    void* hostTables_ = nullptr;
    size_t hostTablesCap_ = 0;
    hipGraphExec_t exec_ = nullptr;
    hipGraph_t graph_ = nullptr;
    // Nodes whose grid is set to the batch's unit count before each launch.
    hipGraphNode_t costNode_ = nullptr;
    hipGraphNode_t wtaNode_[2] = {nullptr, nullptr};
    Tables argTables_{};                 // the tables argument of the graph's kernels
    size_t capCostLds_ = 0, capWta0Lds_ = 0, capWta1Lds_ = 0;
    const void* capPtrs_[4] = {nullptr, nullptr, nullptr, nullptr};
    bool capBig_ = false, capSpeckle_ = false;
};

}  // namespace lumen
