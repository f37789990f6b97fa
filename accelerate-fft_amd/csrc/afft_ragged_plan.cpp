// afft_ragged_plan.cpp -- plans for ragged batches (afft_ragged_* in include/afft.h).
//
// Planning: transforms are grouped by length into classes; each class gets a kernel path
// (direct LDS transform, fused Bluestein, four-step, or four-step Bluestein), its FP64-generated
// tables, and is cut into work items.  A long (four-step) transform becomes one item per column
// tile of each stage, and stage p runs in phase p; the short items fill the phases to balance
// their estimated cost.  Within a phase, items are ordered by decreasing cost
// (longest-processing-time first).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <new>
#include <numeric>
#include <tuple>
#include <vector>

#include "../../include/afft.h"
#include "afft_plan.hpp"
#include "afft_ragged.hpp"

namespace afft {

struct RaggedPlan {
    int dtype = 0, device = 0;
    int64_t count = 0, total = 0;
    Plan tables;                 // owns every device allocation
    RaggedClass *d_classes = nullptr;
    RaggedItem *d_items = nullptr;
    int64_t *d_offsets = nullptr; // member offsets in class order
    int64_t *d_big_off = nullptr;
    void *d_scratch = nullptr;
    bool has_long = false;
    int nitems = 0;
    // phase p runs stage p of every long transform; phases are separate launches, so a stage
    // sees every store of the one before it
    int nphase = 0;
    int phase_off[kStagesBlue] = {}, phase_n[kStagesBlue] = {};
    int64_t phase_grid[kStagesBlue] = {};
    size_t lds = 0;
    std::mutex graph_mu;
    std::map<std::tuple<int, const void *, void *>, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
};

namespace {

constexpr int kMaxE[2] = {16, 10};
// LDS a work item may use (a few workgroups stay resident per CU)
constexpr size_t kLdsBudget[2] = {40 * 1024, 72 * 1024};

size_t esz_of(int dtype) { return dtype == AFFT_C64 ? 8 : 16; }

// direct layout of length n within the ragged kernel's limits, or false
bool ragged_layout(int64_t n, int dtype, LdsLayout &L)
{
    if (!choose_layout(n, dtype, L, kMaxE[dtype], kRaggedBlock)) return false;
    return L.npass <= 1 || (size_t)lds_padded_len((int)n) * esz_of(dtype) <= kLdsBudget[dtype];
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

    bool sub(int64_t len, RaggedSub &S)
    {
        LdsLayout L{};
        if (!ragged_layout(len, rp.dtype, L)) return false;
        const void *tw = nullptr;
        int st = table_pass_twiddles(rp.tables, dev, s, len, L, S.pt, tw);
        if (st) {
            status = st;
            return false;
        }
        S.tw = tw;
        S.E = L.E;
        const int T = (int)(len / L.E);
        S.C = std::max(1, kRaggedBlock / T);
        while (S.C > 1 && (size_t)S.C * lds_padded_len((int)len) * esz_of(rp.dtype) > budget) S.C /= 2;
        lds_need = std::max(lds_need, (size_t)S.C * lds_padded_len((int)len) * esz_of(rp.dtype));
        return true;
    }

    // four-step factors of Lx, both with a direct layout, as balanced as possible
    bool two_phase(int64_t Lx, RaggedClass &K)
    {
        int64_t best1 = 0;
        double best = 1e300;
        for (int64_t d = 2; d < Lx && d <= (1 << 14); ++d) {
            if (Lx % d) continue;
            LdsLayout a{}, b{};
            if (!ragged_layout(d, rp.dtype, a) || !ragged_layout(Lx / d, rp.dtype, b)) continue;
            // balanced factors: measured faster than fewest-pass splits, whose long strided
            // columns coalesce badly
            const double sc = std::fabs(std::log((double)d) - std::log((double)(Lx / d)));
            if (sc < best) {
                best = sc;
                best1 = d;
            }
        }
        if (!best1) return false;
        K.Lx = Lx;
        if (!sub(best1, K.s1) || !sub(Lx / best1, K.s2)) return false;
        int sh = 0;
        while (((int64_t)1 << (2 * sh)) < Lx) ++sh;
        const int64_t lo_cnt = (int64_t)1 << sh, hi_cnt = (Lx + lo_cnt - 1) / lo_cnt;
        int st = table_power(rp.tables, dev, s, Lx, 1, lo_cnt, K.lo);
        if (!st) st = table_power(rp.tables, dev, s, Lx, lo_cnt, hi_cnt, K.hi);
        if (st) {
            status = st;
            return false;
        }
        K.shift = sh;
        return true;
    }

    bool make_class(int64_t L, RaggedClass &K)
    {
        K = RaggedClass{};
        K.L = (int)L;
        LdsLayout lay{};
        const void *tw = nullptr;
        if (is_5smooth(L) && ragged_layout(L, rp.dtype, lay)) {
            K.kind = kDirect;
            K.E = lay.E;
            int st = table_pass_twiddles(rp.tables, dev, s, L, lay, K.pt, tw);
            if (st) {
                status = st;
                return false;
            }
            K.tw = tw;
            K.C = tile_count(L, lay.E, rp.dtype, budget, lay.npass);
            if (lay.npass > 1)
                lds_need = std::max(lds_need, (size_t)K.C * lds_padded_len((int)L) * esz_of(rp.dtype));
            return true;
        }
        if (is_5smooth(L)) {
            K.kind = kTwoPhase;
            return two_phase(L, K);
        }
        const int64_t M = bluestein_length(L);
        K.M = (int)M;
        int st = table_bluestein(rp.tables, dev, s, L, M, K.chirp, K.filt);
        if (st) {
            status = st;
            return false;
        }
        if (ragged_layout(M, rp.dtype, lay)) {
            K.kind = kBlue;
            K.E = lay.E;
            st = table_pass_twiddles(rp.tables, dev, s, M, lay, K.pt, tw);
            if (st) {
                status = st;
                return false;
            }
            K.tw = tw;
            K.C = tile_count(M, lay.E, rp.dtype, budget, lay.npass);
            if (lay.npass > 1)
                lds_need = std::max(lds_need, (size_t)K.C * lds_padded_len((int)M) * esz_of(rp.dtype));
            return true;
        }
        K.kind = kBlueTwo;
        return two_phase(M, K);
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
    // classes
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
    // short transforms: one item per tile of C transforms; long transforms: one item per
    // column tile of every stage, in the phase of that stage
    std::vector<Tmp> shorts, phase[kStagesBlue];
    std::vector<int64_t> big_off;
    int64_t scratch = 0;
    for (auto &kv : by_len) {
        RaggedClass K{};
        if (!b.make_class(kv.first, K)) return b.status ? b.status : AFFT_ERROR_UNSUPPORTED_SIZE;
        const int cls = (int)classes.size();
        classes.push_back(K);
        const int first = (int)members.size();
        members.insert(members.end(), kv.second.begin(), kv.second.end());
        const int cnt = (int)kv.second.size();
        if (K.kind == kDirect || K.kind == kBlue) {
            const double len = (double)(K.kind == kDirect ? K.L : K.M);
            const double per = len * std::log2(std::max(2.0, len)) * (K.kind == kBlue ? 2.5 : 1.0);
            for (int j = 0; j < cnt; j += K.C) {
                const int c = std::min(K.C, cnt - j);
                shorts.push_back({{cls, first + j, c, 0, 0, 0}, per * c});
            }
            continue;
        }
        rp.has_long = true;
        const int nst = K.kind == kBlueTwo ? kStagesBlue : kStagesTwo;
        rp.nphase = std::max(rp.nphase, nst);
        const int64_t L1 = K.s1.pt.n, L2 = K.Lx / L1;
        for (int j = 0; j < cnt; ++j) {
            int big = 0;
            if (K.kind == kBlueTwo) {
                big = (int)big_off.size();
                big_off.push_back(scratch);
                scratch += 2 * K.Lx;
            }
            for (int st = 0; st < nst; ++st) {
                const RaggedSub &S = st % 2 == 0 ? K.s1 : K.s2;
                const int64_t cols = st % 2 == 0 ? L2 : L1;
                const int tiles = (int)((cols + S.C - 1) / S.C);
                const double len = (double)S.pt.n;
                const double per = len * std::log2(std::max(2.0, len)) * S.C * 1.3;
                for (int tl = 0; tl < tiles; ++tl) phase[st].push_back({{cls, first + j, 1, st, tl, big}, per});
            }
        }
    }
    rp.nphase = std::max(rp.nphase, 1);
    // the short items go where the phases are lightest, largest first
    std::stable_sort(shorts.begin(), shorts.end(), [](const Tmp &x, const Tmp &y) { return x.cost > y.cost; });
    double load[kStagesBlue] = {};
    for (int p = 0; p < rp.nphase; ++p)
        for (const Tmp &t : phase[p]) load[p] += t.cost;
    for (const Tmp &t : shorts) {
        int p = 0;
        for (int q = 1; q < rp.nphase; ++q)
            if (load[q] < load[p]) p = q;
        phase[p].push_back(t);
        load[p] += t.cost;
    }
    std::vector<RaggedItem> its;
    for (int p = 0; p < rp.nphase; ++p) {
        std::stable_sort(phase[p].begin(), phase[p].end(), [](const Tmp &x, const Tmp &y) { return x.cost > y.cost; });
        rp.phase_off[p] = (int)its.size();
        rp.phase_n[p] = (int)phase[p].size();
        for (auto &t : phase[p]) its.push_back(t.it);
    }
    rp.nitems = (int)its.size();
    rp.lds = b.lds_need;
    int st = AFFT_SUCCESS;
    rp.d_classes = upload(rp, dev, classes, s, st);
    if (!st) rp.d_items = upload(rp, dev, its, s, st);
    std::vector<int64_t> moffs(members.size());
    for (size_t i = 0; i < members.size(); ++i) moffs[i] = offsets[members[i]];
    if (!st) rp.d_offsets = upload(rp, dev, moffs, s, st);
    if (!st) rp.d_big_off = upload(rp, dev, big_off, s, st);
    if (!st && scratch > 0) rp.d_scratch = plan_alloc(rp.tables, dev, scratch * (int64_t)esz_of(rp.dtype), st);
    if (st) return st;
    const int64_t resident = (int64_t)dev.cus * ragged_occupancy(rp.dtype, rp.lds);
    for (int p = 0; p < rp.nphase; ++p) rp.phase_grid[p] = std::min<int64_t>(resident, rp.phase_n[p]);
    if (hipStreamSynchronize(s) != hipSuccess) return AFFT_ERROR_HIP;
    return AFFT_SUCCESS;
}

static int ragged_launch(const RaggedPlan &rp, int mode, const void *in, void *out, hipStream_t s)
{
    if (rp.nitems == 0) return AFFT_SUCCESS;
    RaggedArgs a{};
    a.in = in;
    a.out = out;
    a.big_off = rp.d_big_off;
    a.classes = rp.d_classes;
    a.moffs = rp.d_offsets;
    a.scratch = rp.d_scratch;
    a.conj = mode == AFFT_FORWARD ? 0 : 1;
    a.inverse = mode == AFFT_INVERSE ? 1 : 0;
    for (int p = 0; p < rp.nphase; ++p) {
        if (rp.phase_n[p] == 0) continue;
        a.items = rp.d_items + rp.phase_off[p];
        a.nitems = rp.phase_n[p];
        if (launch_ragged(rp.dtype, a, rp.phase_grid[p], rp.lds, s) != hipSuccess) return AFFT_ERROR_HIP;
    }
    return AFFT_SUCCESS;
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
    int old = -1;
    (void)hipGetDevice(&old);
    if (hipSetDevice(device) != hipSuccess) return AFFT_ERROR_HIP;
    RaggedPlan *rp = new (std::nothrow) RaggedPlan();
    if (!rp) return AFFT_ERROR_ALLOC;
    rp->dtype = dtype;
    rp->device = device;
    rp->count = count;
    hipStream_t s;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        delete rp;
        return AFFT_ERROR_HIP;
    }
    st = build_ragged(*rp, lengths, dev, s);
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    if (old >= 0) (void)hipSetDevice(old);
    if (st) {
        free_plan(rp->tables);
        delete rp;
        return st;
    }
    *plan = reinterpret_cast<afft_ragged_plan *>(rp);
    return AFFT_SUCCESS;
}

extern "C" int afft_ragged_exec(const afft_ragged_plan *plan, int mode, const void *d_in, void *d_out,
                                void *stream)
{
    if (!plan) return AFFT_ERROR_INVALID_VALUE;
    if (mode != AFFT_FORWARD && mode != AFFT_REVERSE && mode != AFFT_INVERSE) return AFFT_ERROR_INVALID_VALUE;
    const RaggedPlan *rp = reinterpret_cast<const RaggedPlan *>(plan);
    if (rp->total == 0) return AFFT_SUCCESS;
    if (!d_in || !d_out) return AFFT_ERROR_INVALID_VALUE;
    int old = -1;
    (void)hipGetDevice(&old);
    if (old != rp->device && hipSetDevice(rp->device) != hipSuccess) return AFFT_ERROR_HIP;
    hipStream_t s = (hipStream_t)stream;
    const void *src = d_in;
    void *tmp = nullptr;
    // stage A of a long transform writes transposed columns into the output: in place, the
    // input is read from a private copy
    if (d_in == d_out && rp->has_long) {
        const size_t bytes = (size_t)rp->total * (rp->dtype == AFFT_C64 ? 8 : 16);
        if (hipMallocAsync(&tmp, bytes, s) != hipSuccess) return AFFT_ERROR_ALLOC;
        if (hipMemcpyAsync(tmp, d_in, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess) {
            (void)hipFreeAsync(tmp, s);
            return AFFT_ERROR_HIP;
        }
        src = tmp;
    }
    const int st = ragged_launch(*rp, mode, src, d_out, s);
    if (tmp) (void)hipFreeAsync(tmp, s);
    if (old >= 0 && old != rp->device) (void)hipSetDevice(old);
    return st;
}

// Capture the phase launches once per (mode, in, out) on a private stream and replay the
// instantiated graph on the caller's stream (capture is not allowed on the legacy default
// stream).  The graph holds only kernel nodes: the kernel keeps no device-side state, so there
// is no memset node to replay.
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
    int st = AFFT_SUCCESS;
    hipGraphExec_t exec = nullptr;
    {
        std::lock_guard<std::mutex> lk(rp->graph_mu);
        auto key = std::make_tuple(mode, d_in, d_out);
        auto it = rp->graphs.find(key);
        if (it != rp->graphs.end()) {
            exec = it->second.second;
        } else {
            hipStream_t cap;
            if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess) st = AFFT_ERROR_HIP;
            hipGraph_t g = nullptr;
            if (!st && d_in == d_out && rp->has_long) st = AFFT_ERROR_INVALID_VALUE; // needs a copy: use afft_ragged_exec
            if (!st && hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) != hipSuccess) st = AFFT_ERROR_HIP;
            if (!st) {
                const int lst = ragged_launch(*rp, mode, d_in, d_out, cap);
                if (hipStreamEndCapture(cap, &g) != hipSuccess || lst) st = lst ? lst : AFFT_ERROR_HIP;
            }
            if (!st && hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) != hipSuccess) st = AFFT_ERROR_HIP;
            if (!st) rp->graphs[key] = {g, exec};
            else if (g) (void)hipGraphDestroy(g);
            (void)hipStreamDestroy(cap);
        }
    }
    if (!st && hipGraphLaunch(exec, (hipStream_t)stream) != hipSuccess) st = AFFT_ERROR_HIP;
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
    for (auto &kv : rp->graphs) {
        (void)hipGraphExecDestroy(kv.second.second);
        (void)hipGraphDestroy(kv.second.first);
    }
    free_plan(rp->tables);
    if (old >= 0) (void)hipSetDevice(old);
    delete rp;
    return AFFT_SUCCESS;
}

extern "C" int afft_ragged_total(const afft_ragged_plan *plan, int64_t *total)
{
    if (!plan || !total) return AFFT_ERROR_INVALID_VALUE;
    *total = reinterpret_cast<const RaggedPlan *>(plan)->total;
    return AFFT_SUCCESS;
}
