// sgbm_tune.hip - the tuned schedule table: which Schedule a pair of a given width, padded
// disparity range, cost mode and path count runs with on the per-pair path.
//
// The table is a text file written by the autotuner (lumenstereo_amd/tune.py), one entry per
// line:  width_max  Dp  mode  paths  agg_prefetch  agg_priority  census_tile  [agg_wide
// [agg_occupancy]] ('#' comments; a missing agg_wide or agg_occupancy reads as 0).
// A pair uses the entry of its (Dp, mode, paths) with the smallest width_max >= its width (the
// largest one when it is wider than all), and the default schedule when there is none. The
// file next to the library (tuning.tsv) is read on first use; lumen_tuning_load replaces the
// table. A schedule never changes a result, so a missing or stale table only costs speed.
#include <dlfcn.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "lumen_sgbm.h"
#include "sgbm_host.h"

namespace lumen {

namespace {

struct Entry {
    int wmax, dp, mode, paths;
    Schedule s;
};

std::mutex g_mu;
std::vector<Entry> g_table;
bool g_loaded = false;

std::string library_dir() {
    Dl_info info{};
    if (dladdr(reinterpret_cast<const void*>(&library_dir), &info) == 0 || !info.dli_fname)
        return ".";
    std::string p(info.dli_fname);
    const size_t slash = p.rfind('/');
    return slash == std::string::npos ? "." : p.substr(0, slash);
}

bool valid_entry(const Entry& e, std::string& why) {
    // A padded range is one padded_range gives: 96, or a multiple of 64.
    if (e.wmax < 1 || e.dp < 64 || e.dp != padded_range(e.dp) || e.dp > kMaxDisparityRange) {
        why = "bad width or padded range";
        return false;
    }
    if ((e.mode != 0 && e.mode != 1) || (e.paths != 4 && e.paths != 8)) {
        why = "bad cost mode or path count";
        return false;
    }
    if (!valid_schedule(e.s, e.dp / kScanLanes, why)) return false;
    return true;
}

int parse(const std::string& path, std::vector<Entry>& out, std::string& err) {
    std::ifstream f(path);
    if (!f) {
        err = "cannot open " + path;
        return -1;
    }
    std::string line;
    int lineNo = 0;
    while (std::getline(f, line)) {
        ++lineNo;
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::istringstream is(line);
        Entry e{};
        if (!(is >> e.wmax)) continue;  // blank or comment line
        if (!(is >> e.dp >> e.mode >> e.paths >> e.s.aggPrefetch >> e.s.aggPriority >>
              e.s.censusTile)) {
            err = path + ":" + std::to_string(lineNo) + ": expected 7 to 9 integers";
            return -1;
        }
        // Optional trailing columns: agg_wide, then agg_occupancy (0 when absent).
        int* optional[2] = {&e.s.aggWide, &e.s.aggOccupancy};
        for (int* field : optional) {
            *field = 0;
            if (!(is >> *field)) {
                if (!is.eof()) {
                    err = path + ":" + std::to_string(lineNo) + ": expected 7 to 9 integers";
                    return -1;
                }
                *field = 0;
                break;
            }
        }
        std::string extra;
        if (is.clear(), is >> extra) {
            err = path + ":" + std::to_string(lineNo) + ": expected 7 to 9 integers";
            return -1;
        }
        std::string why;
        if (!valid_entry(e, why)) {
            err = path + ":" + std::to_string(lineNo) + ": " + why;
            return -1;
        }
        out.push_back(e);
    }
    return static_cast<int>(out.size());
}

void ensure_loaded_locked() {
    if (g_loaded) return;
    g_loaded = true;
    std::vector<Entry> t;
    std::string err;
    const std::string path = library_dir() + "/tuning.tsv";
    std::ifstream probe(path);
    if (probe && parse(path, t, err) >= 0) g_table = std::move(t);
}

}  // namespace

bool valid_schedule(const Schedule& s, int K, std::string& why) {
    bool ok = true;
    switch (K) {
        case 4: ok = valid_prefetch(4, s.aggPrefetch); break;
        case 6: ok = valid_prefetch(6, s.aggPrefetch); break;
        case 8: ok = valid_prefetch(8, s.aggPrefetch); break;
        case 12: ok = valid_prefetch(12, s.aggPrefetch); break;
        case 16: ok = valid_prefetch(16, s.aggPrefetch); break;
        case 20: ok = valid_prefetch(20, s.aggPrefetch); break;
        case 24: ok = valid_prefetch(24, s.aggPrefetch); break;
        case 28: ok = valid_prefetch(28, s.aggPrefetch); break;
        case 32: ok = valid_prefetch(32, s.aggPrefetch); break;
        default: ok = false;
    }
    if (!ok) {
        why = "aggregation prefetch " + std::to_string(s.aggPrefetch) + " is not supported for " +
              std::to_string(K * kScanLanes) + " padded disparities";
        return false;
    }
    if (s.aggPriority < 0 || s.aggPriority > 2) {
        why = "aggregation priority must be 0, 1 or 2";
        return false;
    }
    if (s.censusTile != 128 && s.censusTile != 256 && s.censusTile != 512) {
        why = "census tile must be 128, 256 or 512";
        return false;
    }
    if (s.aggWide < 0 || s.aggWide > 6) {
        why = "aggregation lane setting must be 0 to 6";
        return false;
    }
    if (s.aggOccupancy != 0 && (s.aggOccupancy < 2 || s.aggOccupancy > 4)) {
        why = "aggregation occupancy must be 0, 2, 3 or 4 blocks per CU";
        return false;
    }
    if (s.aggWide >= 1 && s.aggWide <= 3 && !wide_scanlines_ok(K, s.aggPrefetch)) {
        why = "aggregation prefetch " + std::to_string(s.aggPrefetch) +
              " is not supported on 32-lane scanlines for " + std::to_string(K * kScanLanes) +
              " padded disparities";
        return false;
    }
    if (s.aggWide >= 4 && !xwide_scanlines_ok(K, s.aggPrefetch)) {
        why = K % 8 != 0 ? "64-lane scanlines need a padded range that is a multiple of 128, not " +
                               std::to_string(K * kScanLanes)
                         : "aggregation prefetch " + std::to_string(s.aggPrefetch) +
                               " is not supported on 64-lane scanlines for " +
                               std::to_string(K * kScanLanes) + " padded disparities";
        return false;
    }
    return true;
}

Schedule tuned_schedule(int W, int Dp, int mode, int npaths) {
    std::lock_guard<std::mutex> g(g_mu);
    ensure_loaded_locked();
    const int md = mode == 1 ? 1 : 0;
    const Entry* best = nullptr;
    const Entry* widest = nullptr;
    for (const Entry& e : g_table) {
        if (e.dp != Dp || e.mode != md || e.paths != npaths) continue;
        if (!widest || e.wmax > widest->wmax) widest = &e;
        if (e.wmax >= W && (!best || e.wmax < best->wmax)) best = &e;
    }
    if (!best) best = widest;
    return best ? best->s : default_schedule(Dp / kScanLanes);
}

}  // namespace lumen

extern "C" {

int lumen_tuning_load(const char* path, int* entries_out) {
    std::vector<lumen::Entry> t;
    std::string err;
    const std::string p = (path && *path) ? std::string(path) : lumen::library_dir() + "/tuning.tsv";
    const int n = lumen::parse(p, t, err);
    if (n < 0) {
        lumen::set_last_error(err);
        return LUMEN_EINVAL;
    }
    std::lock_guard<std::mutex> g(lumen::g_mu);
    lumen::g_table = std::move(t);
    lumen::g_loaded = true;
    if (entries_out) *entries_out = n;
    return LUMEN_OK;
}

int lumen_tuning_clear(void) {
    std::lock_guard<std::mutex> g(lumen::g_mu);
    lumen::g_table.clear();
    lumen::g_loaded = true;
    return LUMEN_OK;
}

}  // extern "C"
