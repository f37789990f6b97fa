// afft_tune.cpp -- the autotuner and its persisted table (afft_autotune, afft_ragged_autotune,
// afft_tune_save, afft_tune_load in include/afft.h).
//
// What is tuned: every layout decision the planner makes for a length, which it asks for through
// tune_choice():
//   rows     the kernel variant of the persistent row path (elements per thread, prefetch,
//            occupancy target; see afft_rows.hip)
//   strided  a single-pass strided tile: elements per thread E (which fixes the radices and their
//            order) and columns per workgroup C, encoded C * 1000 + E
//   regcol   lanes per column of the register-resident kernel (1, or 2 / 4 through DPP)
//   ragged   the layout of a class of the ragged persistent kernel: 0 for its compiled layout,
//            else the cap on elements per thread of the pass-table layout (which fixes its radices
//            and the transforms per work item)
//   ragged_split  for a whole ragged batch (keyed by a hash of its length classes): the size in
//            elements below which a long class runs on the side stream
// afft_autotune records the decision points of the plan it would build, then, one point after the
// other, builds the whole plan with each candidate forced (the choices made so far kept), times
// it on the device with HIP events, and records the fastest (the default unless another candidate
// is at least 2% faster, beyond the run-to-run noise); afft_ragged_autotune does the same
// per length class of a ragged batch, timing each class alone (a class of the persistent kernel
// on at least 2^23 elements).  Plans built afterwards use the
// choices.  The table is keyed by (gfx arch, HIP runtime version, kind, dtype, length) and
// persisted as text, one entry per line:
//     afft-tune 1
//     <arch> <hip version> <kind> <dtype> <length> <value> <microseconds>
// Lines for another arch or runtime, malformed lines, unknown kinds and values the planner can
// no longer use are ignored (on load, and again when a plan asks), so a stale or corrupt file
// costs a re-tune, never a wrong plan.
//
// Environment: AFFT_AUTOTUNE=1 tunes a row length the first time a plan needs it;
// AFFT_CACHE_DIR=<dir> loads <dir>/afft_tune.txt at the first plan and saves it after tuning.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/afft.h"
#include "afft_plan.hpp"

namespace afft {

namespace {

const char *const kKindName[kTuneKinds] = {"rows", "strided", "regcol", "ragged", "ragged_split"};

struct TuneEntry {
    int value;
    double us;
};

using TableKey = std::tuple<std::string, int, int, int, int64_t>; // (arch, hipver, kind, dtype, n)
std::mutex g_mu;
std::map<TableKey, TuneEntry> g_table;
bool g_env_loaded = false;

// choices forced on the calling thread while the autotuner times a candidate, and the decision
// points the planner met on it while recording
thread_local std::map<std::tuple<int, int, int64_t>, int> tl_forced;
thread_local std::vector<TunePoint> *tl_record = nullptr;
// set while this thread tunes a row length under g_mu (AFFT_AUTOTUNE=1): the plans it builds then
// take only forced choices and do not look at the table
thread_local bool tl_in_locked_tune = false;

std::string arch_of(int device)
{
    static std::mutex mu;
    static std::map<int, std::string> known;
    std::lock_guard<std::mutex> lk(mu);
    auto it = known.find(device);
    if (it != known.end()) return it->second;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) return "unknown";
    std::string a = p.gcnArchName;
    const size_t colon = a.find(':');
    if (colon != std::string::npos) a = a.substr(0, colon);
    known[device] = a;
    return a;
}

int hip_version()
{
    int v = 0;
    (void)hipRuntimeGetVersion(&v);
    return v;
}

std::string cache_file()
{
    const char *d = std::getenv("AFFT_CACHE_DIR");
    if (!d || !*d) return std::string();
    return std::string(d) + "/afft_tune.txt";
}

int kind_of(const char *name)
{
    for (int k = 0; k < kTuneKinds; ++k)
        if (std::strcmp(name, kKindName[k]) == 0) return k;
    return -1;
}

// values a file may hold for a kind (a plan checks again against its candidates)
bool plausible(int kind, int dtype, int64_t n, int value)
{
    switch (kind) {
    case kTuneRows: return value >= 0 && value < row_variants(dtype, (int)n);
    case kTuneStrided: {
        const int C = value / 1000, E = value % 1000;
        return C >= 1 && C <= 64 && (C & (C - 1)) == 0 && E >= 1 && n % E == 0;
    }
    case kTuneRegcol: return (value == 1 || value == 2 || value == 4) && regcol_supported(dtype, n);
    case kTuneRagged: return value >= 0 && value <= 32; // 0: the compiled layout
    case kTuneRaggedSplit: return value >= 0;          // elements below which a class goes aside
    default: return false;
    }
}

int load_locked(const char *path)
{
    FILE *f = std::fopen(path, "r");
    if (!f) return AFFT_ERROR_IO;
    char line[512];
    if (!std::fgets(line, sizeof line, f) || std::strncmp(line, "afft-tune 1", 11) != 0) {
        std::fclose(f);
        return AFFT_ERROR_IO;
    }
    while (std::fgets(line, sizeof line, f)) {
        char arch[128], kname[32];
        int hv, dtype, value;
        long long n;
        double us;
        if (std::sscanf(line, "%127s %d %31s %d %lld %d %lf", arch, &hv, kname, &dtype, &n, &value, &us) != 7)
            continue; // malformed: ignored
        const int kind = kind_of(kname);
        if (kind < 0 || (dtype != AFFT_C64 && dtype != AFFT_C128) || n < 1 || n >= ((int64_t)1 << 31)) continue;
        if (!plausible(kind, dtype, n, value)) continue; // stale
        g_table[TableKey(std::string(arch), hv, kind, dtype, (int64_t)n)] = {value, us};
    }
    std::fclose(f);
    return AFFT_SUCCESS;
}

int save_locked(const char *path)
{
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "w");
    if (!f) return AFFT_ERROR_IO;
    std::fprintf(f, "afft-tune 1\n");
    for (auto &kv : g_table)
        std::fprintf(f, "%s %d %s %d %lld %d %.3f\n", std::get<0>(kv.first).c_str(), std::get<1>(kv.first),
                     kKindName[std::get<2>(kv.first)], std::get<3>(kv.first), (long long)std::get<4>(kv.first),
                     kv.second.value, kv.second.us);
    const bool ok = std::fclose(f) == 0;
    if (!ok || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return AFFT_ERROR_IO;
    }
    return AFFT_SUCCESS;
}

void ensure_env_loaded_locked()
{
    if (g_env_loaded) return;
    g_env_loaded = true;
    const std::string f = cache_file();
    if (!f.empty()) (void)load_locked(f.c_str());
}

void save_env_locked()
{
    const std::string f = cache_file();
    if (!f.empty()) (void)save_locked(f.c_str());
}

} // namespace

// Device time of one call in microseconds: the median of three batches of five calls after three
// warm-up calls, or a negative value when a call fails.
double tune_time_calls(const std::function<int(hipStream_t)> &call, hipStream_t s)
{
    hipEvent_t a = nullptr, b = nullptr;
    double us = -1;
    if (hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess) {
        bool ok = true;
        for (int i = 0; i < 3 && ok; ++i) ok = call(s) == AFFT_SUCCESS;
        std::vector<double> t;
        for (int rep = 0; rep < 3 && ok; ++rep) {
            ok = hipEventRecord(a, s) == hipSuccess;
            for (int i = 0; i < 5 && ok; ++i) ok = call(s) == AFFT_SUCCESS;
            ok = ok && hipEventRecord(b, s) == hipSuccess && hipEventSynchronize(b) == hipSuccess;
            float ms = 0;
            if (ok && hipEventElapsedTime(&ms, a, b) == hipSuccess) t.push_back(1e3 * ms / 5);
        }
        if (ok && t.size() == 3) {
            std::sort(t.begin(), t.end());
            us = t[1];
        }
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    return us;
}

namespace {

// time one row-plan variant on a (rows x n) batch; microseconds or a negative value
double time_row_variant(int dtype, int64_t n, int variant, int device)
{
    const int64_t rows = std::max<int64_t>(1, ((int64_t)256 << 20) / (n * (dtype == AFFT_C64 ? 8 : 16)));
    const int64_t ext[2] = {rows, n};
    tl_forced[std::make_tuple((int)kTuneRows, dtype, n)] = variant;
    Plan *p = nullptr;
    const int st = create_plan(p, dtype, 2, ext, AFFT_AXES_INNERMOST, device);
    tl_forced.erase(std::make_tuple((int)kTuneRows, dtype, n));
    if (st) return -1;
    const size_t bytes = (size_t)(rows * n) * (dtype == AFFT_C64 ? 8 : 16);
    void *in = nullptr, *out = nullptr;
    double us = -1;
    if (hipMalloc(&in, bytes) == hipSuccess && hipMalloc(&out, bytes) == hipSuccess &&
        hipMemset(in, 0, bytes) == hipSuccess)
        us = tune_time_calls([&](hipStream_t s) { return exec_plan(*p, -1, in, out, 1.0, s); }, nullptr);
    if (in) (void)hipFree(in);
    if (out) (void)hipFree(out);
    free_plan(*p);
    delete p;
    return us;
}

int tune_rows_locked(int dtype, int64_t n, int device, int &best)
{
    const int nv = row_variants(dtype, (int)n);
    best = 0;
    if (nv <= 1) return AFFT_SUCCESS;
    double best_us = 1e300;
    tl_in_locked_tune = true;
    for (int v = 0; v < nv; ++v) {
        const double us = time_row_variant(dtype, n, v, device);
        if (us > 0 && us < best_us) {
            best_us = us;
            best = v;
        }
    }
    tl_in_locked_tune = false;
    if (best_us == 1e300) return AFFT_ERROR_HIP;
    g_table[TableKey(arch_of(device), hip_version(), kTuneRows, dtype, n)] = {best, best_us};
    return AFFT_SUCCESS;
}

} // namespace

int tune_choice(int kind, int dtype, int64_t n, int device, const std::vector<int> &cands)
{
    if (tl_record) {
        const bool seen = std::any_of(tl_record->begin(), tl_record->end(), [&](const TunePoint &p) {
            return p.kind == kind && p.dtype == dtype && p.n == n;
        });
        if (!seen && cands.size() > 1) tl_record->push_back(TunePoint{kind, dtype, n, cands});
    }
    auto f = tl_forced.find(std::make_tuple(kind, dtype, n));
    if (f != tl_forced.end() && std::find(cands.begin(), cands.end(), f->second) != cands.end()) return f->second;
    if (tl_in_locked_tune) return -1;
    std::lock_guard<std::mutex> lk(g_mu);
    ensure_env_loaded_locked();
    auto it = g_table.find(TableKey(arch_of(device), hip_version(), kind, dtype, n));
    if (it != g_table.end() && std::find(cands.begin(), cands.end(), it->second.value) != cands.end())
        return it->second.value;
    if (kind == kTuneRows) {
        const char *at = std::getenv("AFFT_AUTOTUNE");
        if (at && at[0] == '1' && row_variants(dtype, (int)n) > 1) {
            int best = 0;
            if (tune_rows_locked(dtype, n, device, best) == AFFT_SUCCESS) {
                save_env_locked();
                return best;
            }
        }
    }
    return -1;
}

void tune_record(std::vector<TunePoint> *points) { tl_record = points; }
std::vector<TunePoint> *tune_recording() { return tl_record; }

// how much faster than the default a candidate must time to replace it
constexpr double kTuneMargin = 0.02;

int tune_points(const std::vector<TunePoint> &points, int device,
                const std::function<double()> &time_with_forced)
{
    int st = AFFT_SUCCESS;
    for (const TunePoint &p : points) {
        const auto key = std::make_tuple(p.kind, p.dtype, p.n);
        double best_us = 1e300;
        int best = p.cands.front();
        for (int v : p.cands) {
            tl_forced[key] = v;
            const double us = time_with_forced();
            // the default (timed first) yields only to a candidate faster by kTuneMargin: closer
            // timings are within the run-to-run noise, and a noisy pick can cost more elsewhere
            // than it gains here (the ragged kernel's shared layout, the balance of two streams)
            const bool beats = v == p.cands.front() ? us < best_us : us < best_us * (1.0 - kTuneMargin);
            if (us > 0 && beats) {
                best_us = us;
                best = v;
            }
        }
        if (best_us == 1e300) {
            st = AFFT_ERROR_HIP;
            break;
        }
        tl_forced[key] = best; // later points are timed with this choice
        std::lock_guard<std::mutex> lk(g_mu);
        g_table[TableKey(arch_of(device), hip_version(), p.kind, p.dtype, p.n)] = {best, best_us};
    }
    tl_forced.clear();
    std::lock_guard<std::mutex> lk(g_mu);
    if (!st) save_env_locked();
    return st;
}

} // namespace afft

using namespace afft;

extern "C" int afft_autotune(int dtype, int rank, const int64_t *extents, int axes, int device)
{
    if (dtype != AFFT_C64 && dtype != AFFT_C128) return AFFT_ERROR_UNSUPPORTED_DTYPE;
    if (rank < 1 || rank > 8) return AFFT_ERROR_UNSUPPORTED_RANK;
    if (!extents) return AFFT_ERROR_INVALID_VALUE;
    if (axes != AFFT_AXES_INNERMOST && axes != AFFT_AXES_ALL && axes != AFFT_AXES_LAST2)
        return AFFT_ERROR_INVALID_VALUE;
    int64_t total = 1;
    for (int i = 0; i < rank; ++i) {
        if (extents[i] < 0) return AFFT_ERROR_INVALID_VALUE;
        if (__builtin_mul_overflow(total, extents[i], &total)) return AFFT_ERROR_UNSUPPORTED_SIZE;
    }
    if (total == 0) return AFFT_SUCCESS;
    DeviceInfo dev;
    int st = device_info(device, dev);
    if (st) return st;
    int old = -1;
    (void)hipGetDevice(&old);
    if (hipSetDevice(device) != hipSuccess) return AFFT_ERROR_HIP;
    // the decision points of the plan, recorded while it is built once
    std::vector<TunePoint> points;
    {
        Plan *p = nullptr;
        tune_record(&points);
        st = create_plan(p, dtype, rank, extents, axes, device);
        tune_record(nullptr);
        if (!st) {
            free_plan(*p);
            delete p;
        }
    }
    const size_t bytes = (size_t)total * (dtype == AFFT_C64 ? 8 : 16);
    void *in = nullptr, *out = nullptr;
    if (!st && !points.empty()) {
        if (hipMalloc(&in, bytes) != hipSuccess || hipMalloc(&out, bytes) != hipSuccess ||
            hipMemset(in, 0, bytes) != hipSuccess)
            st = AFFT_ERROR_ALLOC;
    }
    if (!st && !points.empty()) {
        st = tune_points(points, device, [&]() -> double {
            Plan *p = nullptr;
            if (create_plan(p, dtype, rank, extents, axes, device)) return -1;
            const double us = tune_time_calls([&](hipStream_t s) { return exec_plan(*p, -1, in, out, 1.0, s); }, nullptr);
            (void)hipDeviceSynchronize();
            free_plan(*p);
            delete p;
            return us;
        });
    }
    if (in) (void)hipFree(in);
    if (out) (void)hipFree(out);
    // plans built from now on pick up the tuned choices
    clear_cache();
    if (old >= 0) (void)hipSetDevice(old);
    return st;
}

extern "C" int afft_tune_save(const char *path)
{
    if (!path) return AFFT_ERROR_INVALID_VALUE;
    std::lock_guard<std::mutex> lk(g_mu);
    return save_locked(path);
}

extern "C" int afft_tune_load(const char *path)
{
    if (!path) return AFFT_ERROR_INVALID_VALUE;
    std::lock_guard<std::mutex> lk(g_mu);
    return load_locked(path);
}
