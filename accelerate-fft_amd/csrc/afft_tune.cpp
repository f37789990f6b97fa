// afft_tune.cpp -- the autotuner and its persisted table (afft_autotune, afft_tune_save,
// afft_tune_load in include/afft.h).
//
// What is tuned: the kernel variant of the persistent row path (elements per thread, prefetch
// on/off, occupancy target; see afft_rows.hip) for a (dtype, length).  afft_autotune builds
// the plan once per candidate, times each on the device with HIP events, and records the
// fastest; plans built afterwards use it.  The table is keyed by (gfx arch, HIP runtime
// version, dtype, length) and persisted as text, one entry per line:
//     afft-tune 1
//     <arch> <hip version> rows <dtype> <length> <variant> <microseconds>
// Lines for another arch or runtime, malformed lines and variants that no longer exist are
// ignored on load, so a stale or corrupt file costs a re-tune, never a wrong plan.
//
// Environment: AFFT_AUTOTUNE=1 tunes a row length the first time a plan needs it;
// AFFT_CACHE_DIR=<dir> loads <dir>/afft_tune.txt at the first plan and saves it after tuning.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/afft.h"
#include "afft_plan.hpp"

namespace afft {

namespace {

struct TuneEntry {
    int variant;
    double us;
};

std::mutex g_mu;
std::map<std::tuple<std::string, int, int, int64_t>, TuneEntry> g_table; // (arch, hipver, dtype, n)
bool g_env_loaded = false;

std::string arch_of(int device)
{
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) return "unknown";
    std::string a = p.gcnArchName;
    const size_t colon = a.find(':');
    return colon == std::string::npos ? a : a.substr(0, colon);
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

int load_locked(const char *path)
{
    FILE *f = std::fopen(path, "r");
    if (!f) return AFFT_ERROR_IO;
    char line[512];
    int status = AFFT_SUCCESS;
    if (!std::fgets(line, sizeof line, f) || std::strncmp(line, "afft-tune 1", 11) != 0) {
        std::fclose(f);
        return AFFT_ERROR_IO;
    }
    while (std::fgets(line, sizeof line, f)) {
        char arch[128], kind[32];
        int hv, dtype, variant;
        long long n;
        double us;
        if (std::sscanf(line, "%127s %d %31s %d %lld %d %lf", arch, &hv, kind, &dtype, &n, &variant, &us) != 7)
            continue; // malformed: ignored
        if (std::strcmp(kind, "rows") != 0 || (dtype != AFFT_C64 && dtype != AFFT_C128) || n < 1) continue;
        if (variant < 0 || variant >= row_variants(dtype, (int)n)) continue; // stale
        g_table[std::make_tuple(std::string(arch), hv, dtype, (int64_t)n)] = {variant, us};
    }
    std::fclose(f);
    return status;
}

int save_locked(const char *path)
{
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "w");
    if (!f) return AFFT_ERROR_IO;
    std::fprintf(f, "afft-tune 1\n");
    for (auto &kv : g_table)
        std::fprintf(f, "%s %d rows %d %lld %d %.3f\n", std::get<0>(kv.first).c_str(), std::get<1>(kv.first),
                     std::get<2>(kv.first), (long long)std::get<3>(kv.first), kv.second.variant, kv.second.us);
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

// time one row-plan variant on a (rows x n) batch; returns microseconds or a negative value
double time_variant(int dtype, int64_t rows, int64_t n, int variant, int device)
{
    force_row_variant(variant);
    Plan *p = nullptr;
    const int64_t ext[2] = {rows, n};
    const int st = create_plan(p, dtype, 2, ext, AFFT_AXES_INNERMOST, device);
    force_row_variant(-1);
    if (st) return -1;
    const size_t bytes = (size_t)(rows * n) * (dtype == AFFT_C64 ? 8 : 16);
    void *in = nullptr, *out = nullptr;
    double us = -1;
    hipEvent_t a = nullptr, b = nullptr;
    if (hipMalloc(&in, bytes) == hipSuccess && hipMalloc(&out, bytes) == hipSuccess &&
        hipMemset(in, 0, bytes) == hipSuccess && hipEventCreate(&a) == hipSuccess &&
        hipEventCreate(&b) == hipSuccess) {
        bool ok = true;
        for (int i = 0; i < 3 && ok; ++i) ok = exec_plan(*p, -1, in, out, 1.0, nullptr) == AFFT_SUCCESS;
        const int reps = 10;
        ok = ok && hipEventRecord(a, nullptr) == hipSuccess;
        for (int i = 0; i < reps && ok; ++i) ok = exec_plan(*p, -1, in, out, 1.0, nullptr) == AFFT_SUCCESS;
        ok = ok && hipEventRecord(b, nullptr) == hipSuccess && hipEventSynchronize(b) == hipSuccess;
        float ms = 0;
        if (ok && hipEventElapsedTime(&ms, a, b) == hipSuccess) us = 1e3 * ms / reps;
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
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
    // a batch of about 256 MiB keeps every variant's persistent grid full
    const int64_t rows = std::max<int64_t>(1, ((int64_t)256 << 20) / (n * (dtype == AFFT_C64 ? 8 : 16)));
    double best_us = 1e300;
    for (int v = 0; v < nv; ++v) {
        const double us = time_variant(dtype, rows, n, v, device);
        if (us > 0 && us < best_us) {
            best_us = us;
            best = v;
        }
    }
    if (best_us == 1e300) return AFFT_ERROR_HIP;
    g_table[std::make_tuple(arch_of(device), hip_version(), dtype, n)] = {best, best_us};
    return AFFT_SUCCESS;
}

} // namespace

// The tuned row variant for (dtype, n) on `device`, tuning first when AFFT_AUTOTUNE=1; -1 when
// there is no entry.
int tuned_row_variant(int dtype, int64_t n, int device)
{
    std::lock_guard<std::mutex> lk(g_mu);
    ensure_env_loaded_locked();
    const auto key = std::make_tuple(arch_of(device), hip_version(), dtype, n);
    auto it = g_table.find(key);
    if (it != g_table.end()) return it->second.variant;
    const char *at = std::getenv("AFFT_AUTOTUNE");
    if (at && at[0] == '1' && row_variants(dtype, (int)n) > 1) {
        int best = 0;
        if (tune_rows_locked(dtype, n, device, best) == AFFT_SUCCESS) {
            const std::string f = cache_file();
            if (!f.empty()) (void)save_locked(f.c_str());
            return best;
        }
    }
    return -1;
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
    DeviceInfo dev;
    int st = device_info(device, dev);
    if (st) return st;
    int old = -1;
    (void)hipGetDevice(&old);
    if (hipSetDevice(device) != hipSuccess) return AFFT_ERROR_HIP;
    // every contiguous-row length the plan would run with the persistent row kernel
    std::vector<int64_t> lens;
    if (extents[rank - 1] > 1) lens.push_back(extents[rank - 1]);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        ensure_env_loaded_locked();
        for (int64_t n : lens) {
            if (row_variants(dtype, (int)n) <= 1) continue;
            int best = 0;
            st = tune_rows_locked(dtype, n, device, best);
            if (st) break;
        }
        const std::string f = cache_file();
        if (!st && !f.empty()) (void)save_locked(f.c_str());
    }
    // plans built from now on pick up the tuned variant
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
