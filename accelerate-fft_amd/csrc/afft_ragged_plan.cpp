// afft_ragged_plan.cpp -- plans for ragged batches (afft_ragged_* in include/afft.h).
//
// Planning: transforms are grouped by length into classes.  A class whose transform fits one
// workgroup of the persistent kernel (a direct LDS transform, or a fused Bluestein transform
// whose padded length fits) gets its FP64-generated tables and is cut into work items of C
// transforms; the items of all such classes are ordered by decreasing estimated cost
// (longest-processing-time first) and run by one persistent launch.  Every longer class runs
// as a plan of the batched engine on the rows of that class, gathered by their offsets: the
// compiled row kernels, the four-step passes and the Bluestein kernels, each launch filling the
// device with hundreds of same-length rows.  Execution is the persistent launch followed by the
// class plans on one stream, and the graph path captures all of them as one hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <tuple>
#include <vector>

#include "../../include/afft.h"
#include "afft_plan.hpp"
#include "afft_ragged.hpp"

namespace afft {

// the rows of one long class, run by a batched plan of shape (count, L)
struct RaggedLong {
    Plan plan;
    int64_t L = 0;
    const int64_t *rows = nullptr; // device: element offsets of the class's rows
    int64_t elems = 0;             // count * L
    bool small = false;            // too few elements to fill the device: runs on the side stream
};

// A long class below this many elements (32 MiB of FP32 data) is latency- or tail-bound: its one
// or two launches hold a few rows per CU.  Those classes run on the side stream beside the large
// ones, and the threshold balances the two streams (measured on the 10,000-transform batch:
// 1M -> 8M moved the 5120...12288 classes to the side stream, 550 -> 465 us; once the large
// classes took fewer passes, 8M -> 4M moved 12288 x 432 back, 392 -> 368 us, while 2.8M, which
// also moves 8192 x 357, gave 400 us).  This is the default; afft_ragged_autotune times every
// partition of a batch's long classes by size and keeps the best for that batch.
constexpr int64_t kSmallLongElems = 1 << 22;

// The per-execution resources of a ragged batch: a side stream forked from and joined back to
// the caller's stream, with its two events, and one device block holding the workspace of every
// long class.  Each execution leases a set (afft_pool.hpp), so two executions of one plan on two
// streams never share a side stream or a workspace.
struct RaggedExec {
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    void *mem = nullptr;
};
using RaggedPool = ExecPool<RaggedExec>;

// At most this many captured graphs per plan (keyed by mode and buffers); the least recently
// replayed one is dropped, after its last replay has finished, to make room for a new key.
constexpr size_t kMaxGraphs = 8;

struct RaggedGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    RaggedPool::Entry *set = nullptr; // the resources captured into the graph
    hipEvent_t launched = nullptr;    // recorded after the last replay
    uint64_t last_use = 0;
};

struct RaggedPlan {
    int dtype = 0, device = 0;
    int64_t count = 0, total = 0;
    Plan tables;                 // owns the persistent kernel's device allocations
    RaggedClass *d_classes = nullptr;
    RaggedItem *d_items = nullptr;
    int64_t *d_offsets = nullptr; // member offsets in class order
    int nitems = 0;
    int64_t grid = 0;
    size_t lds = 0;
    std::vector<std::unique_ptr<RaggedLong>> longs;
    std::vector<int64_t> ws_off;  // offset of each long class's workspace in RaggedExec::mem
    int64_t ws_bytes = 0;
    RaggedPool pool;
    std::mutex graph_mu;
    std::map<std::tuple<int, const void *, void *>, RaggedGraph> graphs;
    uint64_t graph_clock = 0;
};

namespace {

constexpr int kMaxE[2] = {16, 10};
// LDS a work item may use (a few workgroups stay resident per CU)
constexpr size_t kLdsBudget[2] = {40 * 1024, 72 * 1024};

size_t esz_of(int dtype) { return dtype == AFFT_C64 ? 8 : 16; }

// direct layout of length n within the persistent kernel's limits (elements per thread at most
// maxE), or false
bool ragged_layout(int64_t n, int dtype, LdsLayout &L, int maxE)
{
    if (!choose_layout(n, dtype, L, maxE, kRaggedBlock)) return false;
    return L.npass <= 1 || (size_t)lds_padded_len((int)n) * esz_of(dtype) <= kLdsBudget[dtype];
}

// The layout of a class of the persistent kernel: the compiled layout of its transform length
// (Complex Float lengths of the compiled list; `ct` set), else the pass-table layout under the
// default cap on elements per thread; or the tuned (or forced) one of these candidates: 0 for
// the compiled layout, a cap for each pass-table layout that differs.  `L` is the class's true
// length (the autotuner's key), `len` the transform length (L, or M for Bluestein).
bool class_layout(int64_t L, int64_t len, int dtype, int device, LdsLayout &lay, bool &ct)
{
    ct = false;
    const char *no_ct = std::getenv("AFFT_RAGGED_NO_CT");
    const bool has_ct = dtype == AFFT_C64 && ragged_ct_elems((int)len) > 0 && !(no_ct && no_ct[0] == '1');
    const bool has_table = ragged_layout(len, dtype, lay, kMaxE[dtype]);
    if (!has_ct && !has_table) return false;
    std::vector<int> cands;
    if (has_ct) cands.push_back(0);
    if (has_table) {
        cands.push_back(kMaxE[dtype]);
        std::vector<int> seen = {lay.E};
        for (int m : {8, 4, 2}) {
            LdsLayout l{};
            if (m < kMaxE[dtype] && ragged_layout(len, dtype, l, m) &&
                std::find(seen.begin(), seen.end(), l.E) == seen.end()) {
                cands.push_back(m);
                seen.push_back(l.E);
            }
        }
    }
    const int t = tune_choice(kTuneRagged, dtype, L, device, cands);
    const int pick = t >= 0 ? t : cands.front();
    if (pick == 0) {
        ct = true;
        return true;
    }
    if (pick != kMaxE[dtype]) {
        LdsLayout l{};
        if (ragged_layout(len, dtype, l, pick)) lay = l;
    }
    return true;
}

// transforms per item for a direct/Bluestein class: fill the workgroup, keep LDS in budget
int tile_count(int64_t len, int E, int dtype, size_t budget, int npass)
{
    const int T = (int)(len / E);
    int C = std::max(1, kRaggedBlock / T);
    if (npass > 1)
        while (C > 1 && (size_t)C * lds_padded_len((int)len) * esz_of(dtype) > budget) C /= 2;
    return C;
}

struct Builder {
    RaggedPlan &rp;
    const DeviceInfo &dev;
    hipStream_t s;
    int status = AFFT_SUCCESS;
    size_t budget;
    size_t lds_need = 0;

    // a class of the persistent kernel, or false when the length needs a batched plan
    bool make_class(int64_t L, RaggedClass &K)
    {
        K = RaggedClass{};
        K.L = (int)L;
        LdsLayout lay{};
        const void *tw = nullptr;
        int64_t len = L;
        bool ct = false;
        if (is_5smooth(L)) {
            if (!class_layout(L, L, rp.dtype, rp.device, lay, ct)) return false;
            K.kind = kDirect;
        } else {
            const int64_t M = bluestein_length(L);
            if (!class_layout(L, M, rp.dtype, rp.device, lay, ct)) return false;
            K.kind = kBlue;
            K.M = (int)M;
            len = M;
            int st = table_bluestein(rp.tables, dev, s, L, M, K.chirp, K.filt);
            if (st) {
                status = st;
                return false;
            }
        }
        if (ct) {
            // compiled layout: whole transforms fill the workgroup, twiddles from W_len^m
            K.ctn = (int)len;
            K.E = ragged_ct_elems((int)len);
            const int T = (int)len / K.E;
            K.C = kRaggedBlock / T;
            int st = table_power(rp.tables, dev, s, len, 1, len, tw);
            if (st) {
                status = st;
                return false;
            }
            K.tw = tw;
            if (T > 1) lds_need = std::max(lds_need, (size_t)K.C * lds_padded_len((int)len) * esz_of(rp.dtype));
            return true;
        }
        K.E = lay.E;
        int st = table_pass_twiddles(rp.tables, dev, s, len, lay, K.pt, tw);
        if (st) {
            status = st;
            return false;
        }
        K.tw = tw;
        K.C = tile_count(len, lay.E, rp.dtype, budget, lay.npass);
        if (lay.npass > 1) lds_need = std::max(lds_need, (size_t)K.C * lds_padded_len((int)len) * esz_of(rp.dtype));
        return true;
    }
};

template <typename T> T *upload(RaggedPlan &rp, const DeviceInfo &dev, const std::vector<T> &v, hipStream_t s, int &st)
{
    T *p = (T *)plan_alloc(rp.tables, dev, (int64_t)std::max<size_t>(1, v.size()) * sizeof(T), st);
    if (!p) return nullptr;
    if (!v.empty() && hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess)
        st = AFFT_ERROR_HIP;
    return p;
}

} // namespace

// A key for a ragged batch's length classes (FNV-1a over (length, count) in length order), the
// autotuner's `length` for batch-wide decisions; 62 bits, so it stays a positive int64
static int64_t ragged_batch_key(const std::map<int64_t, std::vector<int>> &by_len)
{
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](uint64_t v) {
        for (int b = 0; b < 8; ++b) {
            h ^= (v >> (8 * b)) & 0xff;
            h *= 1099511628211ull;
        }
    };
    for (const auto &kv : by_len) {
        mix((uint64_t)kv.first);
        mix((uint64_t)kv.second.size());
    }
    return (int64_t)(h >> 2);
}

static int build_ragged(RaggedPlan &rp, const int64_t *lengths, const DeviceInfo &dev, hipStream_t s)
{
    rp.tables.dtype = rp.dtype;
    rp.tables.device = rp.device;
    std::vector<int64_t> offsets(rp.count);
    int64_t acc = 0;
    for (int64_t i = 0; i < rp.count; ++i) {
        offsets[i] = acc;
        acc += lengths[i];
    }
    rp.total = acc;
    std::map<int64_t, std::vector<int>> by_len;
    for (int64_t i = 0; i < rp.count; ++i)
        if (lengths[i] > 0) by_len[lengths[i]].push_back((int)i);
    Builder b{rp, dev, s};
    b.budget = kLdsBudget[rp.dtype];
    std::vector<RaggedClass> classes;
    std::vector<int> members;
    struct Tmp {
        RaggedItem it;
        double cost;
    };
    std::vector<Tmp> items;
    std::vector<int> long_first; // first member of each long class
    for (auto &kv : by_len) {
        const int first = (int)members.size();
        members.insert(members.end(), kv.second.begin(), kv.second.end());
        const int cnt = (int)kv.second.size();
        RaggedClass K{};
        if (!b.make_class(kv.first, K)) {
            if (b.status) return b.status;
            auto lp = std::make_unique<RaggedLong>();
            lp->L = kv.first;
            lp->elems = (int64_t)cnt * kv.first;
            const int64_t ext[2] = {cnt, kv.first};
            // rows packed back to back start at arbitrary 8- or 16-byte offsets
            lp->plan.flags = kPlanUnalignedRows;
            const int st = build_plan(lp->plan, rp.dtype, 2, ext, AFFT_AXES_INNERMOST, rp.device, dev, s);
            rp.longs.push_back(std::move(lp));
            if (st) return st;
            long_first.push_back(first);
            continue;
        }
        const int cls = (int)classes.size();
        classes.push_back(K);
        const double len = (double)(K.kind == kDirect ? K.L : K.M);
        const double per = len * std::log2(std::max(2.0, len)) * (K.kind == kBlue ? 2.5 : 1.0);
        for (int j = 0; j < cnt; j += K.C) {
            const int c = std::min(K.C, cnt - j);
            items.push_back({{cls, first + j, c}, per * c});
        }
    }
    // which long classes go to the side stream: those below a size in elements, the default or
    // the one tuned for this batch (its classes hashed), or AFFT_RAGGED_SMALL
    {
        const char *env = std::getenv("AFFT_RAGGED_SMALL");
        int64_t small_elems = env ? (int64_t)std::atoll(env) : kSmallLongElems;
        if (!env && !rp.longs.empty()) {
            std::vector<int64_t> sizes;
            for (const auto &lp : rp.longs) sizes.push_back(lp->elems);
            std::sort(sizes.begin(), sizes.end());
            // the default, then every other partition: none aside, the smallest 1, 2, ... aside
            auto aside = [&](int64_t t) { return std::lower_bound(sizes.begin(), sizes.end(), t) - sizes.begin(); };
            std::vector<int> cands = {(int)std::min<int64_t>(kSmallLongElems, INT32_MAX)};
            std::vector<int64_t> seen = {aside(cands[0])};
            for (size_t k = 0; k <= sizes.size(); ++k) {
                const int64_t t = std::min<int64_t>(k ? sizes[k - 1] + 1 : 0, INT32_MAX);
                if (std::find(seen.begin(), seen.end(), aside(t)) != seen.end()) continue;
                seen.push_back(aside(t));
                cands.push_back((int)t);
            }
            const int t = tune_choice(kTuneRaggedSplit, rp.dtype, ragged_batch_key(by_len), rp.device, cands);
            if (t >= 0) small_elems = t;
        }
        for (auto &lp : rp.longs) lp->small = lp->elems < small_elems;
    }
    std::stable_sort(items.begin(), items.end(), [](const Tmp &x, const Tmp &y) { return x.cost > y.cost; });
    std::vector<RaggedItem> its;
    for (auto &t : items) its.push_back(t.it);
    rp.nitems = (int)its.size();
    rp.lds = b.lds_need;
    int st = AFFT_SUCCESS;
    rp.d_classes = upload(rp, dev, classes, s, st);
    if (!st) rp.d_items = upload(rp, dev, its, s, st);
    std::vector<int64_t> moffs(members.size());
    for (size_t i = 0; i < members.size(); ++i) moffs[i] = offsets[members[i]];
    if (!st) rp.d_offsets = upload(rp, dev, moffs, s, st);
    if (st) return st;
    for (size_t i = 0; i < rp.longs.size(); ++i) rp.longs[i]->rows = rp.d_offsets + long_first[i];
    for (const auto &lp : rp.longs) {
        rp.ws_off.push_back(rp.ws_bytes);
        rp.ws_bytes += workspace_bytes(lp->plan);
    }
    const int64_t resident = (int64_t)dev.cus * ragged_occupancy(rp.dtype, rp.lds);
    // up to four times the resident workgroups: one item each unless the batch has more items,
    // so that a workgroup retires once its stores are issued and a new one's loads overlap them
    // (one resident round walking the list: 4096 x 4096 137 us, the 10,000-transform batch's
    // persistent classes 67 us and the 2^27 batch 965 us untuned, against 143 / 65 / 921 us)
    rp.grid = std::max<int64_t>(1, std::min<int64_t>(4 * resident, rp.nitems));
    if (hipStreamSynchronize(s) != hipSuccess) return AFFT_ERROR_HIP;
    return AFFT_SUCCESS;
}

static int make_exec(const RaggedPlan &rp, RaggedExec &x)
{
    x = RaggedExec{};
    if (hipStreamCreateWithFlags(&x.side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&x.fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&x.join, hipEventDisableTiming) != hipSuccess)
        return AFFT_ERROR_HIP;
    if (rp.ws_bytes > 0 && device_alloc(&x.mem, (size_t)rp.ws_bytes) != hipSuccess) return AFFT_ERROR_ALLOC;
    return AFFT_SUCCESS;
}

static void free_exec(RaggedExec &x)
{
    if (x.side) {
        (void)hipStreamSynchronize(x.side);
        (void)hipStreamDestroy(x.side);
    }
    if (x.fork) (void)hipEventDestroy(x.fork);
    if (x.join) (void)hipEventDestroy(x.join);
    if (x.mem) (void)hipFree(x.mem);
    x = RaggedExec{};
}

static void destroy_ragged(RaggedPlan *rp)
{
    for (auto &kv : rp->graphs) {
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
        if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
        if (kv.second.launched) (void)hipEventDestroy(kv.second.launched);
    }
    rp->graphs.clear();
    rp->pool.clear(free_exec);
    for (auto &lp : rp->longs) free_plan(lp->plan);
    free_plan(rp->tables);
    delete rp;
}

static int run_long(const RaggedLong &lp, int mode, const void *in, void *out, hipStream_t s, const Workspace &ws)
{
    const int direction = mode == AFFT_FORWARD ? -1 : +1;
    const double scale = mode == AFFT_INVERSE ? 1.0 / (double)lp.L : 1.0;
    return exec_plan(lp.plan, direction, in, out, scale, s, lp.rows, lp.rows, nullptr, nullptr, &ws);
}

// The persistent launch and the small long classes on the set's side stream, the other long
// classes on s, joined back into s.  Every transform reads and writes only its own rows, so the
// branches touch disjoint memory and in-place execution (in == out) needs no copy.
static int ragged_launch(const RaggedPlan &rp, int mode, const void *in, void *out, hipStream_t s,
                         const RaggedExec &x)
{
    static const bool serial = [] {
        const char *v = std::getenv("AFFT_RAGGED_SERIAL");
        return v && v[0] == '1';
    }();
    hipStream_t side = serial ? s : x.side;
    if (hipEventRecord(x.fork, s) != hipSuccess || hipStreamWaitEvent(side, x.fork, 0) != hipSuccess)
        return AFFT_ERROR_HIP;
    auto ws_of = [&](size_t i) {
        return carve_workspace(rp.longs[i]->plan, x.mem ? static_cast<char *>(x.mem) + rp.ws_off[i] : nullptr);
    };
    int st = AFFT_SUCCESS;
    // the bandwidth-bound long classes on the caller's stream first: they are the critical path,
    // and enqueued first they start without waiting for the side stream's launches to be issued
    for (size_t i = 0; i < rp.longs.size(); ++i)
        if (!st && !rp.longs[i]->small) st = run_long(*rp.longs[i], mode, in, out, s, ws_of(i));
    if (!st && rp.nitems > 0) {
        RaggedArgs a{};
        a.in = in;
        a.out = out;
        a.nitems = rp.nitems;
        a.classes = rp.d_classes;
        a.items = rp.d_items;
        a.moffs = rp.d_offsets;
        a.conj = mode == AFFT_FORWARD ? 0 : 1;
        a.inverse = mode == AFFT_INVERSE ? 1 : 0;
        if (launch_ragged(rp.dtype, a, rp.grid, rp.lds, side) != hipSuccess) st = AFFT_ERROR_HIP;
    }
    for (size_t i = 0; i < rp.longs.size(); ++i)
        if (!st && rp.longs[i]->small) st = run_long(*rp.longs[i], mode, in, out, side, ws_of(i));
    // join even after an error, so that the side stream never runs ahead of the caller's
    if (hipEventRecord(x.join, side) != hipSuccess || hipStreamWaitEvent(s, x.join, 0) != hipSuccess)
        st = st ? st : AFFT_ERROR_HIP;
    return st;
}

} // namespace afft

using namespace afft;

extern "C" int afft_ragged_plan_create(afft_ragged_plan **plan, int dtype, int64_t count,
                                       const int64_t *lengths, int device)
{
    if (!plan) return AFFT_ERROR_INVALID_VALUE;
    *plan = nullptr;
    if (dtype != AFFT_C64 && dtype != AFFT_C128) return AFFT_ERROR_UNSUPPORTED_DTYPE;
    if (count < 0 || (count > 0 && !lengths) || count > (int64_t)1 << 30) return AFFT_ERROR_INVALID_VALUE;
    for (int64_t i = 0; i < count; ++i)
        if (lengths[i] < 0 || lengths[i] > ((int64_t)1 << 30)) return AFFT_ERROR_INVALID_VALUE;
    DeviceInfo dev;
    int st = device_info(device, dev);
    if (st) return st;
    reset_alloc_count();
    int old = -1;
    (void)hipGetDevice(&old);
    if (hipSetDevice(device) != hipSuccess) return AFFT_ERROR_HIP;
    RaggedPlan *rp = new (std::nothrow) RaggedPlan();
    hipStream_t s = nullptr;
    if (!rp || hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        const int err = rp ? AFFT_ERROR_HIP : AFFT_ERROR_ALLOC;
        delete rp;
        if (old >= 0) (void)hipSetDevice(old);
        return err;
    }
    rp->dtype = dtype;
    rp->device = device;
    rp->count = count;
    st = build_ragged(*rp, lengths, dev, s);
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    if (st) destroy_ragged(rp);
    if (old >= 0) (void)hipSetDevice(old);
    if (st) return st;
    *plan = reinterpret_cast<afft_ragged_plan *>(rp);
    return AFFT_SUCCESS;
}

extern "C" int afft_ragged_exec(const afft_ragged_plan *plan, int mode, const void *d_in, void *d_out,
                                void *stream)
{
    if (!plan) return AFFT_ERROR_INVALID_VALUE;
    if (mode != AFFT_FORWARD && mode != AFFT_REVERSE && mode != AFFT_INVERSE) return AFFT_ERROR_INVALID_VALUE;
    RaggedPlan *rp = const_cast<RaggedPlan *>(reinterpret_cast<const RaggedPlan *>(plan));
    if (rp->total == 0) return AFFT_SUCCESS;
    if (!d_in || !d_out) return AFFT_ERROR_INVALID_VALUE;
    int old = -1;
    (void)hipGetDevice(&old);
    if (old != rp->device && hipSetDevice(rp->device) != hipSuccess) return AFFT_ERROR_HIP;
    hipStream_t s = (hipStream_t)stream;
    auto make = [rp](RaggedExec &x) { return make_exec(*rp, x); };
    RaggedPool::Entry *e = nullptr;
    int st = stream_capturing(s) ? rp->pool.acquire_pinned(make, e) : rp->pool.acquire(s, make, e);
    if (!st) {
        st = ragged_launch(*rp, mode, d_in, d_out, s, e->res);
        rp->pool.release(e, s, true);
    }
    if (old >= 0 && old != rp->device) (void)hipSetDevice(old);
    return st;
}

// Capture the launches once per (mode, in, out) on a private stream, with a resource set of the
// graph's own, and replay the instantiated graph on the caller's stream (capture is not allowed on
// the legacy default stream).  The graph holds only kernel nodes and the fork/join of its own side
// stream: no kernel keeps device-side state, so there is no memset node to replay.  Replays of one
// graph are ordered through its `launched` event, since they share the graph's workspace.
extern "C" int afft_ragged_exec_graph(afft_ragged_plan *plan, int mode, const void *d_in, void *d_out,
                                      void *stream)
{
    if (!plan) return AFFT_ERROR_INVALID_VALUE;
    if (mode != AFFT_FORWARD && mode != AFFT_REVERSE && mode != AFFT_INVERSE) return AFFT_ERROR_INVALID_VALUE;
    RaggedPlan *rp = reinterpret_cast<RaggedPlan *>(plan);
    if (rp->total == 0) return AFFT_SUCCESS;
    if (!d_in || !d_out) return AFFT_ERROR_INVALID_VALUE;
    int old = -1;
    (void)hipGetDevice(&old);
    if (old != rp->device && hipSetDevice(rp->device) != hipSuccess) return AFFT_ERROR_HIP;
    hipStream_t s = (hipStream_t)stream;
    int st = AFFT_SUCCESS;
    {
        std::lock_guard<std::mutex> lk(rp->graph_mu);
        auto key = std::make_tuple(mode, d_in, d_out);
        auto it = rp->graphs.find(key);
        if (it == rp->graphs.end()) {
            if (rp->graphs.size() >= kMaxGraphs) {
                auto lru = rp->graphs.begin();
                for (auto j = rp->graphs.begin(); j != rp->graphs.end(); ++j)
                    if (j->second.last_use < lru->second.last_use) lru = j;
                RaggedGraph &g = lru->second;
                (void)hipEventSynchronize(g.launched);
                (void)hipGraphExecDestroy(g.exec);
                (void)hipGraphDestroy(g.graph);
                (void)hipEventDestroy(g.launched);
                rp->pool.drop(g.set, free_exec);
                rp->graphs.erase(lru);
            }
            RaggedGraph g;
            auto make = [rp](RaggedExec &x) { return make_exec(*rp, x); };
            st = rp->pool.acquire_pinned(make, g.set);
            hipStream_t cap = nullptr;
            if (!st && hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess) st = AFFT_ERROR_HIP;
            if (!st && hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) != hipSuccess) st = AFFT_ERROR_HIP;
            if (!st) {
                const int lst = ragged_launch(*rp, mode, d_in, d_out, cap, g.set->res);
                if (hipStreamEndCapture(cap, &g.graph) != hipSuccess || lst) st = lst ? lst : AFFT_ERROR_HIP;
            }
            if (!st && hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0) != hipSuccess) st = AFFT_ERROR_HIP;
            if (!st && hipEventCreateWithFlags(&g.launched, hipEventDisableTiming) != hipSuccess) st = AFFT_ERROR_HIP;
            if (cap) (void)hipStreamDestroy(cap);
            if (st) {
                if (g.exec) (void)hipGraphExecDestroy(g.exec);
                if (g.graph) (void)hipGraphDestroy(g.graph);
                if (g.set) rp->pool.drop(g.set, free_exec);
            } else {
                it = rp->graphs.emplace(key, g).first;
            }
        }
        if (!st) {
            RaggedGraph &g = it->second;
            g.last_use = ++rp->graph_clock;
            if (hipStreamWaitEvent(s, g.launched, 0) != hipSuccess || hipGraphLaunch(g.exec, s) != hipSuccess ||
                hipEventRecord(g.launched, s) != hipSuccess)
                st = AFFT_ERROR_HIP;
        }
    }
    if (old >= 0 && old != rp->device) (void)hipSetDevice(old);
    return st;
}

extern "C" int afft_ragged_plan_destroy(afft_ragged_plan *plan)
{
    if (!plan) return AFFT_ERROR_INVALID_VALUE;
    RaggedPlan *rp = reinterpret_cast<RaggedPlan *>(plan);
    int old = -1;
    (void)hipGetDevice(&old);
    (void)hipSetDevice(rp->device);
    (void)hipDeviceSynchronize();
    destroy_ragged(rp);
    if (old >= 0) (void)hipSetDevice(old);
    return AFFT_SUCCESS;
}

// Tune every length class of a ragged batch on its own: the class's transforms as a ragged plan
// of one length, its decision points (the persistent kernel's layout, or the layouts inside the
// batched plan of a long class) timed candidate by candidate.  A class of the persistent kernel
// is timed on at least 2^23 elements (its length repeated): a few hundred short transforms take
// less device time than the host takes to issue the call, which would hide the layout.  Then,
// with those layouts, the whole batch: which of its long classes run on the side stream.
extern "C" int afft_ragged_autotune(int dtype, int64_t count, const int64_t *lengths, int device)
{
    if (dtype != AFFT_C64 && dtype != AFFT_C128) return AFFT_ERROR_UNSUPPORTED_DTYPE;
    if (count < 0 || (count > 0 && !lengths) || count > (int64_t)1 << 30) return AFFT_ERROR_INVALID_VALUE;
    std::map<int64_t, int64_t> by_len;
    for (int64_t i = 0; i < count; ++i) {
        if (lengths[i] < 0 || lengths[i] > ((int64_t)1 << 30)) return AFFT_ERROR_INVALID_VALUE;
        if (lengths[i] > 0) ++by_len[lengths[i]];
    }
    int old = -1;
    (void)hipGetDevice(&old);
    if (hipSetDevice(device) != hipSuccess) return AFFT_ERROR_HIP;
    const size_t esz = dtype == AFFT_C64 ? 8 : 16;
    int st = AFFT_SUCCESS;
    for (auto &kv : by_len) {
        const std::vector<int64_t> one((size_t)kv.second, kv.first);
        std::vector<TunePoint> points;
        afft_ragged_plan *p = nullptr;
        tune_record(&points);
        st = afft_ragged_plan_create(&p, dtype, (int64_t)one.size(), one.data(), device);
        tune_record(nullptr);
        if (st) break;
        (void)afft_ragged_plan_destroy(p);
        points.erase(std::remove_if(points.begin(), points.end(),
                                    [](const TunePoint &q) { return q.kind == kTuneRaggedSplit; }),
                     points.end());
        if (points.empty()) continue;
        const bool persistent = std::all_of(points.begin(), points.end(),
                                            [](const TunePoint &q) { return q.kind == kTuneRagged; });
        const int64_t reps = persistent ? std::max<int64_t>(kv.second, ((int64_t)1 << 23) / kv.first) : kv.second;
        const std::vector<int64_t> timed((size_t)reps, kv.first);
        void *buf = nullptr;
        const size_t bytes = (size_t)(kv.first * reps) * esz;
        if (hipMalloc(&buf, bytes) != hipSuccess || hipMemset(buf, 0, bytes) != hipSuccess) {
            if (buf) (void)hipFree(buf);
            st = AFFT_ERROR_ALLOC;
            break;
        }
        st = tune_points(points, device, [&]() -> double {
            afft_ragged_plan *q = nullptr;
            if (afft_ragged_plan_create(&q, dtype, (int64_t)timed.size(), timed.data(), device)) return -1;
            const double us = tune_time_calls(
                [&](hipStream_t s) { return afft_ragged_exec(q, AFFT_FORWARD, buf, buf, s); }, nullptr);
            (void)afft_ragged_plan_destroy(q);
            return us;
        });
        (void)hipFree(buf);
        if (st) break;
    }
    if (!st && count > 0) {
        std::vector<TunePoint> points;
        afft_ragged_plan *p = nullptr;
        tune_record(&points);
        st = afft_ragged_plan_create(&p, dtype, count, lengths, device);
        tune_record(nullptr);
        if (!st) (void)afft_ragged_plan_destroy(p);
        points.erase(std::remove_if(points.begin(), points.end(),
                                    [](const TunePoint &q) { return q.kind != kTuneRaggedSplit; }),
                     points.end());
        if (!st && !points.empty()) {
            int64_t total = 0;
            for (auto &kv : by_len) total += kv.first * kv.second;
            // timed as replays of the batch's graph (the form whose launch gaps are smallest, so
            // the balance of the two streams shows most), on a stream of its own
            void *buf = nullptr;
            hipStream_t ts = nullptr;
            const size_t bytes = (size_t)total * esz;
            if (hipMalloc(&buf, bytes) != hipSuccess || hipMemset(buf, 0, bytes) != hipSuccess) {
                st = AFFT_ERROR_ALLOC;
            } else if (hipStreamCreate(&ts) != hipSuccess) {
                st = AFFT_ERROR_HIP;
            } else {
                st = tune_points(points, device, [&]() -> double {
                    afft_ragged_plan *q = nullptr;
                    if (afft_ragged_plan_create(&q, dtype, count, lengths, device)) return -1;
                    const double us = tune_time_calls(
                        [&](hipStream_t s) { return afft_ragged_exec_graph(q, AFFT_FORWARD, buf, buf, s); }, ts);
                    (void)afft_ragged_plan_destroy(q);
                    return us;
                });
            }
            if (ts) (void)hipStreamDestroy(ts);
            if (buf) (void)hipFree(buf);
        }
    }
    if (old >= 0) (void)hipSetDevice(old);
    return st;
}

extern "C" int afft_ragged_total(const afft_ragged_plan *plan, int64_t *total)
{
    if (!plan || !total) return AFFT_ERROR_INVALID_VALUE;
    *total = reinterpret_cast<const RaggedPlan *>(plan)->total;
    return AFFT_SUCCESS;
}
