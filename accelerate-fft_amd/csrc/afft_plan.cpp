// afft_plan.cpp -- planning, the plan cache and the C ABI of the FFT engine.
//
// A plan is an immutable list of kernel launches plus the device tables they read (pass
// twiddles, four-step twiddles, Bluestein chirps and filters, all generated on the device in
// FP64 and rounded to the element type).  It holds no stream and no mutable state, so several
// host threads may execute one plan on different streams at once; this is the fix for the
// shared-handle setStream/exec race the cuFFT plugin leaves open (LLVM/PTX.hs:121-124,
// PTX/Plans.hs:58-63).  The cache is keyed on the full (device, dtype, rank, axes, extents)
// tuple rather than a hash of it (PTX/Plans.hs:72 keys on (context, hash)).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/afft.h"
#include "afft_internal.hpp"
#include "afft_plan.hpp"

namespace afft {

// ----------------------------------------------------------------------------------------------
// Factorisation (host only).

static const int kRadices[] = {16, 8, 5, 4, 3, 2};


// Device allocations of the library go through device_alloc.  Failure injection for tests:
// with AFFT_TEST_FAIL_ALLOC=k set, the k-th allocation (k >= 1) of the calling thread since the
// last AllocCount reset (the start of a plan creation) fails as hipMalloc's would.
static thread_local int64_t tl_alloc_count = 0;
void reset_alloc_count() { tl_alloc_count = 0; }

hipError_t device_alloc(void **p, size_t bytes)
{
    *p = nullptr;
    ++tl_alloc_count;
    if (const char *v = std::getenv("AFFT_TEST_FAIL_ALLOC")) {
        if (std::atoll(v) == tl_alloc_count) return hipErrorOutOfMemory;
    }
    return hipMalloc(p, bytes);
}

// the buffer a four-step split moves its data to: output <-> work, aux <-> aux2
static Slot other(Slot s)
{
    switch (s) {
    case Slot::kOut: return Slot::kWork;
    case Slot::kWork: return Slot::kOut;
    case Slot::kAux: return Slot::kAux2;
    default: return Slot::kAux;
    }
}

static bool getenv_flag(const char *name)
{
    const char *v = std::getenv(name);
    return v && v[0] && v[0] != '0';
}

bool is_5smooth(int64_t n)
{
    if (n < 1) return false;
    for (int p : {2, 3, 5})
        while (n % p == 0) n /= p;
    return n == 1;
}

static int prime_mask(int64_t n)
{
    int m = 0;
    if (n % 2 == 0) m |= 1;
    if (n % 3 == 0) m |= 2;
    if (n % 5 == 0) m |= 4;
    return m;
}

int64_t lds_limit_elems(int dtype, size_t lds_bytes)
{
    // largest n with lds_padded_len(n) elements fitting the LDS budget, rounded down to a
    // 5-smooth-friendly bound: 16384 (c64) / 8192 (c128) on a 160 KiB LDS.
    const size_t esz = dtype == AFFT_C64 ? 8 : 16;
    int64_t n = 1;
    while ((size_t)lds_padded_len((int)(n * 2)) * esz <= lds_bytes) n *= 2;
    return n;
}

bool choose_layout(int64_t n, int dtype, LdsLayout &out, int maxE, int maxT)
{
    if (n < 1 || n > (1 << 30)) return false;
    if (n == 1) {
        out.E = 1;
        out.npass = 0;
        return true;
    }
    if (!is_5smooth(n)) return false;
    const int target = 16;
    bool found = false;
    LdsLayout best{};
    int best_score = 1 << 30;
    const int need = prime_mask(n);
    for (int ie = 0; ie < kNumElemsPerThread; ++ie) {
        const int E = kElemsPerThread[ie];
        if (E > n || n % E != 0) continue;
        if ((prime_mask(E) & need) != need) continue;
        const int64_t T = n / E;
        if (T > max_block_threads(dtype == AFFT_C128, E)) continue;
        if (E > maxE || (maxT > 0 && T > maxT)) continue;
        if (dtype == AFFT_C128 && E > 32) continue; // FP64 data beyond 128 VGPRs
        LdsLayout L{};
        L.E = E;
        int64_t rem = n;
        int np = 0;
        bool ok = true;
        while (rem > 1) {
            int pick = 0;
            for (int R : kRadices)
                if (E % R == 0 && rem % R == 0) {
                    pick = R;
                    break;
                }
            if (!pick || np >= kMaxPasses) {
                ok = false;
                break;
            }
            L.radix[np++] = pick;
            rem /= pick;
        }
        if (!ok) continue;
        L.npass = np;
        const int score = np * 1000 + std::abs(E - target);
        if (score < best_score) {
            best_score = score;
            best = L;
            found = true;
        }
    }
    if (found) out = best;
    return found;
}

// Bluestein padded length: the smallest M >= 2n - 1 of the form 2^k, 3*2^k or 5*2^k.
int64_t bluestein_length(int64_t n)
{
    const int64_t need = 2 * n - 1;
    int64_t best = 0;
    for (int64_t base : {1, 3, 5}) {
        int64_t m = base;
        while (m < need) m *= 2;
        if (best == 0 || m < best) best = m;
    }
    return best;
}

static PassTable make_pass_table(int64_t n, const LdsLayout &L, int64_t *tw_elems)
{
    PassTable pt{};
    pt.n = (int)n;
    pt.npass = L.npass;
    int64_t off = 0;
    int64_t Ns = 1;
    for (int p = 0; p < L.npass; ++p) {
        pt.radix[p] = L.radix[p];
        pt.tw_off[p] = (int)off;
        off += Ns * L.radix[p];
        Ns *= L.radix[p];
    }
    *tw_elems = off;
    return pt;
}

// ----------------------------------------------------------------------------------------------
// Plan construction.

namespace {

struct Builder {
    Plan &plan;
    const DeviceInfo &dev;
    hipStream_t s;
    int status = AFFT_SUCCESS;

    size_t esz() const { return plan.dtype == AFFT_C64 ? 8 : 16; }

    void *alloc(int64_t bytes)
    {
        if (bytes <= 0) return nullptr;
        void *p = nullptr;
        if (device_alloc(&p, (size_t)bytes) != hipSuccess) {
            status = AFFT_ERROR_ALLOC;
            return nullptr;
        }
        plan.allocs.push_back(p);
        plan.device_bytes += bytes;
        return p;
    }

    bool check(hipError_t e)
    {
        if (e != hipSuccess && status == AFFT_SUCCESS) status = AFFT_ERROR_HIP;
        return e == hipSuccess;
    }

    // pass twiddles for length n with layout L (cached per length inside the plan)
    const void *twiddles(int64_t n, const LdsLayout &L, PassTable &pt)
    {
        int64_t cnt = 0;
        pt = make_pass_table(n, L, &cnt);
        auto it = plan.tw_cache.find(n);
        if (it != plan.tw_cache.end()) return it->second;
        void *buf = alloc(std::max<int64_t>(cnt, 1) * (int64_t)esz());
        if (!buf) return nullptr;
        if (cnt > 0) check(gen_pass_twiddles(plan.dtype, buf, pt, s));
        plan.tw_cache[n] = buf;
        return buf;
    }

    // Lanes per transform of the register kernel: AFFT_REGCOL_Q (1, 2 or 4), else the tuned or
    // forced choice, else one.
    int regcol_lanes(int dtype, int64_t n)
    {
        if (const char *v = std::getenv("AFFT_REGCOL_Q")) {
            const int q = std::atoi(v);
            if (q == 1 || q == 2 || q == 4) return q;
        }
        const int q = tune_choice(kTuneRegcol, dtype, n, plan.device, {1, 2, 4});
        return q > 0 ? q : 1;
    }

    // A single-pass strided tile of length n: the tuned or forced (columns, elements per thread),
    // encoded C * 1000 + E, among the layouts of each elements-per-thread cap and the default,
    // half and double column counts that fit the workgroup and the LDS.  Returns false to keep
    // the default layout.
    bool tuned_strided(int64_t n, int64_t ntrans, LdsLayout &L, int &C, int &block, size_t &lds)
    {
        std::vector<int> cands;
        auto add = [&](const LdsLayout &l, int c) {
            const int T = (int)(n / l.E);
            if (c < 1 || (int64_t)c > std::max<int64_t>(1, ntrans)) return;
            if ((int64_t)c * T > max_block_threads(plan.dtype == AFFT_C128, l.E)) return;
            if (l.npass > 1 && (size_t)c * lds_padded_len((int)n) * esz() > dev.lds_bytes) return;
            const int v = c * 1000 + l.E;
            if (std::find(cands.begin(), cands.end(), v) == cands.end()) cands.push_back(v);
        };
        int c0, b0;
        size_t l0;
        shape(n, L.E, true, L.npass, ntrans, c0, b0, l0);
        add(L, c0); // the default first
        for (int maxE : {4, 8, 16, 32}) {
            LdsLayout l{};
            if (!choose_layout(n, plan.dtype, l, maxE)) continue;
            int c, b;
            size_t ld;
            shape(n, l.E, true, l.npass, ntrans, c, b, ld);
            for (int cc : {c, c / 2, c * 2}) add(l, cc);
        }
        const int v = tune_choice(kTuneStrided, plan.dtype, n, plan.device, cands);
        if (v < 0 || v == cands.front()) return false;
        LdsLayout l{};
        if (!choose_layout(n, plan.dtype, l, v % 1000) || l.E != v % 1000) return false;
        L = l;
        C = v / 1000;
        block = C * (int)(n / l.E);
        lds = l.npass > 1 ? (size_t)C * lds_padded_len((int)n) * esz() : 0;
        return true;
    }

    // Columns a strided workgroup aims for: neighbouring lanes take neighbouring columns, so C
    // columns make one C * element-size byte segment per row; at least 16 columns (a 128-byte
    // line in FP32), more while the workgroup stays at or under 256 threads.
    static int strided_cols_target(int T)
    {
        int C = 16;
        while (C < 64 && 2 * C * T <= 256) C *= 2;
        return C;
    }

    // A strided axis whose single-pass tile would read row segments shorter than 128 bytes is
    // split into two passes of balanced length (four-step) whose tiles read whole lines.
    bool split_strided(int64_t n, int64_t stride)
    {
        if (getenv_flag("AFFT_NO_COL_SPLIT")) return false;
        LdsLayout L{};
        if (!choose_layout(n, plan.dtype, L)) return false;
        int C, block;
        size_t lds;
        shape(n, L.E, true, L.npass, stride, C, block, lds);
        return (int64_t)C * (int64_t)esz() < 128 && stride * (int64_t)esz() >= 128;
    }

    // workgroup shape for a transform of length n with T threads, strided or not
    void shape(int64_t n, int E, bool strided, int npass, int64_t ntrans, int &C, int &block, size_t &lds)
    {
        const int T = (int)(n / E);
        const int maxb = max_block_threads(plan.dtype == AFFT_C128, E);
        const size_t per = (size_t)lds_padded_len((int)n) * esz();
        if (!strided) {
            C = std::max(1, std::min(256, maxb) / T);
            if (npass > 1)
                while (C > 1 && (size_t)C * per > 64 * 1024) C /= 2;
        } else {
            C = strided_cols_target(T);
            while (C > 1 && (size_t)C * T > maxb) C /= 2;
            if (npass > 1) {
                while (C > 1 && (size_t)C * per > dev.lds_bytes) C /= 2;
                // beyond whole 128-byte lines (16 columns in FP32, 8 in FP64), keep room for three
                // workgroups per CU
                const int line_cols = (int)(128 / esz());
                while (C > line_cols && (size_t)C * per > dev.lds_bytes / 3) C /= 2;
            }
            // prefer at least 64 threads per workgroup when the LDS allows it
        }
        if ((int64_t)C > ntrans && ntrans > 0) C = (int)std::max<int64_t>(1, ntrans);
        block = C * T;
        lds = npass > 1 ? (size_t)C * per : 0;
    }

    // Transforms j and j + 1 start at neighbouring elements for runs of at least 16, so a wave
    // with one transform per lane reads whole lines.
    static bool lanes_adjacent(const Addr &a)
    {
        return (a.cnt0 >= 16 && a.st0 == 1) || (a.cnt0 == 1 && a.cnt1 >= 16 && a.st1 == 1);
    }

    // Plain single-workgroup transform (n 5-smooth and within the LDS limit).
    void add_stockham(int64_t n, int64_t ntrans, const Addr &ain, const Addr &aout, bool strided,
                      Slot src, Slot dst, int64_t post_L = 0, int post_level = 0,
                      const void *post_hi = nullptr, const void *post_lo = nullptr, int post_shift = 0,
                      const void *post_tab = nullptr, int64_t post_ld = 0)
    {
        if (strided && regcol_supported(plan.dtype, n) && lanes_adjacent(ain) && lanes_adjacent(aout) &&
            !getenv_flag("AFFT_NO_REGCOL")) {
            // one thread per transform, whole in registers (afft_regcol)
            Launch l{};
            l.kind = Launch::kRegCol;
            l.src = src;
            l.dst = dst;
            FftArgs &a = l.fa;
            a.ntrans = ntrans;
            a.ain = ain;
            a.aout = aout;
            a.pt.n = (int)n;
            a.E = (int)n;
            a.C = 1;
            a.xcd_remap = !getenv_flag("AFFT_NO_XCD") ? 1 : 0;
            a.scale = 1.0;
            a.post_L = post_L;
            a.post_level = post_level;
            a.post_hi = post_hi;
            a.post_lo = post_lo;
            a.post_shift = post_shift;
            a.lanes = regcol_lanes(plan.dtype, n);
            l.block = regcol_block();
            l.lds = 0;
            l.grid = (ntrans * a.lanes + l.block - 1) / l.block;
            plan.launches.push_back(l);
            return;
        }
        LdsLayout L{};
        if (!choose_layout(n, plan.dtype, L)) {
            status = AFFT_ERROR_UNSUPPORTED_SIZE;
            return;
        }
        int tC = 0, tblock = 0;
        size_t tlds = 0;
        const bool tuned = strided && tuned_strided(n, ntrans, L, tC, tblock, tlds);
        PassTable pt{};
        const void *tw = twiddles(n, L, pt);
        if (status) return;
        Launch l{};
        l.kind = Launch::kStockham;
        l.src = src;
        l.dst = dst;
        FftArgs &a = l.fa;
        a.tw = tw;
        a.ntrans = ntrans;
        a.ain = ain;
        a.aout = aout;
        a.pt = pt;
        a.E = L.E;
        if (tuned) {
            a.C = tC;
            l.block = tblock;
            l.lds = tlds;
        } else {
            shape(n, L.E, strided, L.npass, ntrans, a.C, l.block, l.lds);
        }
        a.lane_col = strided ? 1 : 0;
        a.lds_interleave = strided ? 1 : 0;
        a.xcd_remap = strided && !getenv_flag("AFFT_NO_XCD") ? 1 : 0;
        a.scale = 1.0;
        a.post_L = post_L;
        a.post_level = post_level;
        a.post_hi = post_hi;
        a.post_lo = post_lo;
        a.post_shift = post_shift;
        a.post_tab = post_tab;
        a.post_ld = post_ld;
        // strided tiles of 128 ... 1024 (Complex Float) and 256 (Complex Double) run compile-time
        // passes (the launcher picks them when the power table is present)
        const bool ct = plan.dtype == AFFT_C64 ? (n == 128 || n == 256 || n == 512 || n == 1024) : n == 256;
        if (strided && ct && L.E == 16 && !getenv_flag("AFFT_NO_CT_COLS")) {
            a.pw = power_table(n);
            if (status) return;
        }
        l.grid = (ntrans + a.C - 1) / a.C;
        plan.launches.push_back(l);
    }

    // Chirp c[m] (m < n) and filter FFT_M(b)/M for a Bluestein length, cached per n.  The
    // filter is transformed in FP64 by a sub-plan of this engine and then rounded.
    bool bluestein_tables(int64_t n, int64_t M, const void *&chirp, const void *&filt)
    {
        auto it = plan.blue_cache.find(n);
        if (it != plan.blue_cache.end()) {
            chirp = it->second.first;
            filt = it->second.second;
            return true;
        }
        void *cbuf = alloc(n * (int64_t)esz());
        void *fbuf = alloc(M * (int64_t)esz());
        if (status) return false;
        check(gen_chirp(plan.dtype, cbuf, n, s));
        void *b64 = nullptr;
        if (device_alloc(&b64, (size_t)M * 16) != hipSuccess) {
            status = AFFT_ERROR_ALLOC;
            return false;
        }
        check(gen_bluestein_kernel_f64(b64, n, M, s));
        Plan sub;
        int64_t ext[1] = {M};
        // the filter's FP64 transform runs once, at planning time: its layout is not tuned
        std::vector<TunePoint> *rec = tune_recording();
        tune_record(nullptr);
        int st = build_plan(sub, AFFT_C128, 1, ext, AFFT_AXES_INNERMOST, plan.device, dev, s);
        tune_record(rec);
        if (st == AFFT_SUCCESS) st = exec_plan(sub, -1, b64, b64, 1.0, s);
        if (st != AFFT_SUCCESS && status == AFFT_SUCCESS) status = st;
        check(convert_scale_from_f64(plan.dtype, fbuf, b64, M, 1.0 / (double)M, s));
        check(hipStreamSynchronize(s));
        (void)hipFree(b64);
        free_plan(sub);
        if (status) return false;
        plan.blue_cache[n] = {cbuf, fbuf};
        chirp = cbuf;
        filt = fbuf;
        return true;
    }

    // Bluestein over several launches when M exceeds one workgroup: pre-multiply into the
    // auxiliary buffer, FFT_M, filter, FFT_M, post-multiply.
    void add_bluestein_large(int64_t n, int64_t ntrans, const Addr &ain, const Addr &aout, Slot src,
                             Slot dst)
    {
        const int64_t M = bluestein_length(n);
        const void *chirp = nullptr, *filt = nullptr;
        if (!bluestein_tables(n, M, chirp, filt)) return;
        plan.aux_elems = std::max(plan.aux_elems, ntrans * M);
        Launch l{};
        l.kind = Launch::kChirp;
        l.ca.chirp = chirp;
        l.ca.filt = filt;
        l.ca.ntrans = ntrans;
        l.ca.n = n;
        l.ca.M = M;
        l.ca.scale = 1.0;
        Launch pre = l, mid = l, post = l;
        pre.src = src;
        pre.dst = Slot::kAux;
        pre.ca.a = ain;
        pre.ca.mode = 0;
        mid.src = mid.dst = Slot::kAux;
        mid.ca.mode = 1;
        post.src = Slot::kAux;
        post.dst = dst;
        post.ca.a = aout;
        post.ca.mode = 2;
        plan.launches.push_back(pre);
        const Slot r1 = add_axis(M, 1, ntrans, Slot::kAux, Slot::kAux);
        if (status) return;
        mid.src = mid.dst = r1;
        plan.launches.push_back(mid);
        const Slot r2 = add_axis(M, 1, ntrans, r1, r1);
        if (status) return;
        post.src = r2;
        plan.launches.push_back(post);
    }

    void add_bluestein(int64_t n, int64_t ntrans, const Addr &ain, const Addr &aout, bool strided,
                       Slot src, Slot dst)
    {
        const int64_t M = bluestein_length(n);
        LdsLayout L{};
        if (M > dev.max_lds_len[plan.dtype] || !choose_layout(M, plan.dtype, L)) {
            add_bluestein_large(n, ntrans, ain, aout, src, dst);
            return;
        }
        PassTable pt{};
        const void *tw = twiddles(M, L, pt);
        if (status) return;
        const void *chirp = nullptr, *filt = nullptr;
        if (!bluestein_tables(n, M, chirp, filt)) return;
        Launch l{};
        l.kind = Launch::kBluestein;
        l.src = src;
        l.dst = dst;
        BluesteinArgs &a = l.ba;
        a.tw = tw;
        a.chirp = chirp;
        a.filt = filt;
        a.ntrans = ntrans;
        a.n = n;
        a.ain = ain;
        a.aout = aout;
        a.pt = pt;
        a.E = L.E;
        int C;
        shape(M, L.E, strided, L.npass, ntrans, C, l.block, l.lds);
        a.C = C;
        a.lane_col = strided ? 1 : 0;
        a.lds_interleave = strided ? 1 : 0;
        a.xcd_remap = strided && !getenv_flag("AFFT_NO_XCD") ? 1 : 0;
        a.scale = 1.0;
        // Complex Float padded lengths 640 and 10240 run compile-time transforms.  Contiguous
        // rows of 10240 take 512 threads a transform (E = 20): 1.3-1.4x faster than 256 threads
        // (E = 40) from 137 to 8192 transforms of 4099 points
        if (plan.dtype == AFFT_C64 && (M == 640 || M == 10240) && !getenv_flag("AFFT_NO_CT_COLS")) {
            a.pw = power_table(M);
            if (status) return;
            if (M == 10240 && !strided) {
                a.E = 20;
                shape(M, 20, strided, 7, ntrans, C, l.block, l.lds);
                a.C = C;
            }
        }
        l.grid = (ntrans + C - 1) / C;
        plan.launches.push_back(l);
    }

    // W_n^m table, m in [0, n) (cached per n inside the plan)
    const void *power_table(int64_t n)
    {
        auto it = plan.pow_cache.find(n);
        if (it != plan.pow_cache.end()) return it->second;
        void *buf = alloc(n * (int64_t)esz());
        if (!buf) return nullptr;
        check(gen_power_table(plan.dtype, buf, n, 1, n, s));
        plan.pow_cache[n] = buf;
        return buf;
    }

    // Persistent row kernel (compiled layout for this n).
    void add_rows(int64_t n, int64_t nrows, Slot src, Slot dst, int variant)
    {
        const void *tw = power_table(n);
        if (status) return;
        Launch l{};
        l.kind = Launch::kRows;
        l.src = src;
        l.dst = dst;
        l.row_n = (int)n;
        l.row_variant = variant;
        l.ra.tw = tw;
        l.ra.nrows = nrows;
        l.ra.scale = 1.0;
        const int per_tile = row_rows_per_tile(plan.dtype, (int)n, variant);
        const int64_t tiles = (nrows + per_tile - 1) / per_tile;
        const int64_t resident = (int64_t)dev.cus * row_occupancy(plan.dtype, (int)n, variant);
        // persistent (one resident round walking the tiles, each workgroup prefetching its next
        // tile) for small batches; from 256 MiB one tile per workgroup, where the dispatcher
        // overlaps a retiring workgroup's stores with the next one's loads; from 128 MiB four
        // rounds for rows up to 2048 points.  Measured: 16384 x 4096 185.8 -> 176.7 us, 1 GiB
        // batches of 256 ... 4096 points 10% faster; the spectral step's 128 MiB rows of 512
        // (forward 162 -> 154 us); the 128 MiB rows of fft2D 4096^2 lost 5 us with four rounds and
        // stay persistent; at 64 MiB the persistent grid is fastest (4096: 21.7 against 26.1 us)
        int64_t rounds = 1;
        const int64_t bytes = nrows * n * (int64_t)esz();
        if (row_grid_beyond_resident(plan.dtype, (int)n, variant)) {
            if (bytes >= ((int64_t)256 << 20)) rounds = tiles;
            else if (bytes >= ((int64_t)128 << 20) && n <= 2048) rounds = 4;
        }
        l.grid = std::max<int64_t>(1, std::min<int64_t>(tiles, resident * std::min<int64_t>(rounds, tiles)));
        l.block = row_block(plan.dtype, (int)n, variant);
        l.lds = row_lds_bytes(plan.dtype, (int)n, variant);
        plan.launches.push_back(l);
    }

    // an explicit AFFT_ROW_VARIANT, else the variant forced by the autotuner or tuned, else the
    // default layout
    int pick_row_variant(int64_t n, int64_t nrows)
    {
        const int nv = row_variants(plan.dtype, (int)n);
        int variant = -1;
        if (const char *v = std::getenv("AFFT_ROW_VARIANT")) {
            variant = std::atoi(v);
        } else {
            // the default first: the autotuner keeps it unless another is measurably faster
            const int def = row_default_variant(plan.dtype, (int)n, nrows * n * (int64_t)esz());
            std::vector<int> cands = {def};
            for (int v = 0; v < nv; ++v)
                if (v != def) cands.push_back(v);
            variant = tune_choice(kTuneRows, plan.dtype, n, plan.device, cands);
            if (variant < 0) variant = def;
        }
        if (variant < 0 || variant >= nv) variant = 0;
        return variant;
    }

    // every matrix (rows x cols, contiguous, back to back) transposed
    void add_transpose(int64_t nbatch, int64_t rows, int64_t cols, Slot src, Slot dst)
    {
        TransposeArgs t{};
        t.rows = rows;
        t.cols = cols;
        t.nbatch = nbatch;
        t.nb1 = 1;
        t.in_bs2 = t.out_bs2 = rows * cols;
        t.in_rs = cols;
        t.out_rs = rows;
        add_transpose(t, src, dst);
    }

    void add_transpose(const TransposeArgs &t, Slot src, Slot dst)
    {
        Launch l{};
        l.kind = Launch::kTranspose;
        l.src = src;
        l.dst = dst;
        l.ta = t;
        l.ta.scale = 1.0;
        l.ta.xcd_remap = !getenv_flag("AFFT_NO_XCD") ? 1 : 0;
        plan.launches.push_back(l);
    }

    // Six-step factors of a contiguous row of length n (n = n1 * n2): a register-resident column
    // length n1 whose rows n2 have a compiled row kernel; 0 if none (then the split six-step
    // below, else the balanced four-step; a six-step with an LDS column pass measured slower than
    // both).
    int64_t six_step_n1(int64_t n)
    {
        if (getenv_flag("AFFT_NO_SIX_STEP")) return 0;
        // (split-exchange rows hold one row per workgroup: too few rows in flight for the
        // middle pass of a six-step)
        for (int64_t c : {64, 32, 16})
            if (n % c == 0 && regcol_supported(plan.dtype, c) && n / c <= (1 << 30) &&
                row_variants(plan.dtype, (int)(n / c)) > 0 && !row_split(plan.dtype, (int)(n / c)))
                return c;
        return 0;
    }

    // Transform along one axis: extent n, element stride s_ (product of the inner extents),
    // o outer blocks; reads slot src, writes slot dst.
    enum class AxisKind { kRows, kSingle, kBluestein, kSplit };

    // How one axis of length n and element stride s_ is transformed, o outer blocks; build_plan
    // and add_axis both follow this decision.  Rows beyond 16384 points with a split-exchange
    // layout run one row per workgroup: a single pass that beats the two-pass six-step on a few
    // hundred rows (20480 x 447: 59 against 63 us) but not on many (20480 x 1638: 203 against
    // 188 us), so they take it for up to 4 rows per CU.
    AxisKind axis_kind(int64_t n, int64_t s_, int64_t o)
    {
        const bool strided = s_ > 1;
        if (!strided && row_variants(plan.dtype, (int)n) > 0 && !getenv_flag("AFFT_NO_ROWS") &&
            !(row_split(plan.dtype, (int)n) && n > 16384 && o > 4 * (int64_t)dev.cus))
            return AxisKind::kRows;
        if (n <= dev.max_lds_len[plan.dtype] && is_5smooth(n) && !(strided && split_strided(n, s_)))
            return AxisKind::kSingle;
        if (!is_5smooth(n)) return AxisKind::kBluestein;
        return AxisKind::kSplit;
    }

    // two-level table for W_n^e: lo[e & (2^sh - 1)], hi[e >> sh]
    bool power_tables(int64_t n, const void *&hi, const void *&lo, int &sh)
    {
        sh = 0;
        while (((int64_t)1 << (2 * sh)) < n) ++sh;
        const int64_t lo_cnt = (int64_t)1 << sh;
        const int64_t hi_cnt = (n + lo_cnt - 1) / lo_cnt;
        void *l = alloc(lo_cnt * (int64_t)esz());
        void *h = alloc(hi_cnt * (int64_t)esz());
        if (status) return false;
        check(gen_power_table(plan.dtype, l, n, 1, lo_cnt, s));
        check(gen_power_table(plan.dtype, h, n, lo_cnt, hi_cnt, s));
        lo = l;
        hi = h;
        return status == AFFT_SUCCESS;
    }

    // Six-step transform of o contiguous rows of length n = n1 * n2 (element i = i1*n2 + i2,
    // output k = k1 + n1*k2), three passes that each read and write whole lines:
    //   A  columns: length n1 over i1 (stride n2), twiddled by W_n^{i2 k1}, in place (src -> cur)
    //   B  rows: length n2 over i2 for every (o, k1), in place on cur (compiled row kernel)
    //   C  transpose of every n1 x n2 matrix into other(cur), putting X[k1 + n1*k2] in place
    void add_six_step(int64_t n, int64_t n1, int64_t o, Slot src, Slot cur)
    {
        const int64_t n2 = n / n1;
        const void *hi = nullptr, *lo = nullptr;
        int sh = 0;
        if (!power_tables(n, hi, lo, sh)) return;
        const Slot next = other(cur);
        int64_t &room = next == Slot::kWork ? plan.work_elems
                        : next == Slot::kAux2 ? plan.aux2_elems : plan.aux_elems;
        if (next != Slot::kOut) room = std::max(room, o * n);
        Addr ia{};
        ia.cnt0 = 1;
        ia.cnt1 = n2;
        ia.st0 = 1;
        ia.st1 = 1;
        ia.st2 = n;
        ia.es = n2;
        add_stockham(n1, o * n2, ia, ia, true, src, cur, n, 1, hi, lo, sh);
        if (status) return;
        // rows short enough for a tile of several whole rows store transposed themselves: two
        // passes instead of three.  A tile's stores are runs of C neighbouring k1 (one 128-byte
        // line or more); its rows are rotated by the output's alignment (RowArgs::tr_rot), so the
        // runs stay whole lines also for transforms that start at arbitrary offsets, as the rows
        // of a ragged batch do (unrotated, their runs straddle two lines shared with the next
        // tile: 40960 x 295 took 111 us with transposed rows against 80 us for rows plus transpose)
        const int tr_rows = row_rows_per_tile(plan.dtype, (int)n2, kRowTransposed);
        if (tr_rows > 1 && n1 % tr_rows == 0 && !getenv_flag("AFFT_NO_ROWS") && !getenv_flag("AFFT_NO_ROW_TR")) {
            add_rows(n2, o * n1, cur, next, kRowTransposed);
            plan.launches.back().ra.tr_n1 = n1;
            plan.launches.back().ra.tr_rot = 1;
            return;
        }
        if (row_variants(plan.dtype, (int)n2) > 0 && !getenv_flag("AFFT_NO_ROWS")) {
            add_rows(n2, o * n1, cur, cur, pick_row_variant(n2, o * n1));
        } else {
            Addr ra{};
            ra.cnt0 = 1;
            ra.cnt1 = 1;
            ra.st2 = n2;
            ra.es = 1;
            add_stockham(n2, o * n1, ra, ra, false, cur, cur);
        }
        if (status) return;
        add_transpose(o, n1, n2, cur, next);
    }

    // Factors of o contiguous rows n = n1 * m1 * m2 for a six-step whose rows n2 = m1 * m2 have
    // no compiled row kernel: a register column length n1, a strided single-pass length m1 whose
    // tiles read whole lines, and a row length m2 with a transposed-store form.  The register
    // pass runs one thread per column, o * n / n1 threads, and hides its load latency only with
    // about 8 waves per CU: the longest n1 that keeps that many, else the shortest.  False if none.
    bool split_six_step_factors(int64_t n, int64_t o, int64_t &n1, int64_t &m1, int64_t &m2)
    {
        if ((plan.flags & kPlanUnalignedRows) || getenv_flag("AFFT_NO_SPLIT_SIX") || getenv_flag("AFFT_NO_ROWS") ||
            getenv_flag("AFFT_NO_ROW_TR"))
            return false;
        const int64_t min_threads = (int64_t)dev.cus * 8 * 64;
        bool found = false;
        for (int64_t c : {64, 32, 16}) {
            if (n % c || !regcol_supported(plan.dtype, c)) continue;
            const int64_t n2 = n / c;
            for (int64_t r = 1024; r >= 16; --r) {
                if (n2 % r) continue;
                const int tr = row_rows_per_tile(plan.dtype, (int)r, kRowTransposed);
                const int64_t q = n2 / r;
                if (tr <= 1 || q < 2 || (c * q) % tr) continue;
                if (q > dev.max_lds_len[plan.dtype] || !is_5smooth(q) || split_strided(q, r)) continue;
                n1 = c;
                m1 = q;
                m2 = r;
                found = true;
                break;
            }
            if (found && o * n / n1 >= min_threads) return true;
        }
        return found;
    }

    // Six-step transform of o contiguous rows n = n1 * n2 whose rows n2 = m1 * m2 are themselves
    // split (element i = i1*n2 + j1*m2 + j2, output k = k1 + n1*(q1 + m1*q2)); three passes that
    // each read and write whole lines:
    //   A  columns: length n1 over i1 (stride n2), twiddled by W_n^{i2 k1}, in place (src -> cur)
    //   B  columns: length m1 over j1 (stride m2) for every (o, k1, j2), twiddled by
    //      W_n2^{j2 q1}, in place on cur
    //   C  rows: length m2 over j2 for every (o, q1, k1), taken k1 fastest through a row table and
    //      stored transposed: X[k1 + n1*q1 + n1*m1*q2] into other(cur)
    void add_split_six_step(int64_t n, int64_t n1, int64_t m1, int64_t m2, int64_t o, Slot src, Slot cur)
    {
        const int64_t n2 = m1 * m2;
        const void *hi = nullptr, *lo = nullptr, *hi2 = nullptr, *lo2 = nullptr;
        int sh = 0, sh2 = 0;
        if (!power_tables(n, hi, lo, sh) || !power_tables(n2, hi2, lo2, sh2)) return;
        const Slot next = other(cur);
        int64_t &room = next == Slot::kWork ? plan.work_elems
                        : next == Slot::kAux2 ? plan.aux2_elems : plan.aux_elems;
        if (next != Slot::kOut) room = std::max(room, o * n);
        Addr ia{};
        ia.cnt0 = 1;
        ia.cnt1 = n2;
        ia.st0 = 1;
        ia.st1 = 1;
        ia.st2 = n;
        ia.es = n2;
        add_stockham(n1, o * n2, ia, ia, true, src, cur, n, 1, hi, lo, sh);
        if (status) return;
        Addr ja{};
        ja.cnt0 = m2;
        ja.cnt1 = n1;
        ja.st0 = 1;
        ja.st1 = n2;
        ja.st2 = n;
        ja.es = m2;
        // the twiddles W_n2^{j2 q1} as one m1 x m2 grid (n2 elements): per element one load,
        // coalesced across a tile's lanes, where the two-level table took two scattered loads
        // (measured: those lookups cost 5 of the 47 us of a 2^22 row)
        void *grid = alloc(n2 * (int64_t)esz());
        if (status) return;
        check(gen_twiddle_grid(plan.dtype, grid, n2, m1, m2, s));
        if (status) return;
        add_stockham(m1, o * n1 * m2, ja, ja, true, cur, cur, n2, 0, hi2, lo2, sh2, grid, m2);
        if (status) return;
        // row r = k1 + n1*(q1 + m1*b) of transform b starts at b*n + k1*n2 + q1*m2
        const int64_t nrows = o * n1 * m1;
        std::vector<int64_t> offs((size_t)nrows);
        for (int64_t r = 0; r < nrows; ++r) {
            const int64_t k1 = r % n1, q1 = (r / n1) % m1, b = r / (n1 * m1);
            offs[(size_t)r] = b * n + k1 * n2 + q1 * m2;
        }
        void *table = alloc(nrows * (int64_t)sizeof(int64_t));
        if (status) return;
        check(hipMemcpyAsync(table, offs.data(), (size_t)nrows * sizeof(int64_t), hipMemcpyHostToDevice, s));
        check(hipStreamSynchronize(s));
        if (status) return;
        add_rows(m2, nrows, cur, next, kRowTransposed);
        plan.launches.back().ra.tr_n1 = n1 * m1;
        plan.launches.back().ra.roffs = static_cast<const int64_t *>(table);
    }

    // Transform one axis, reading slot src.  Single-pass kinds run in place or out of place and
    // leave the result in `cur`; a four-step split leaves it in other(cur), since its second pass
    // permutes the elements and cannot run in place.  Returns the slot holding the result.
    Slot add_axis(int64_t n, int64_t s_, int64_t o, Slot src, Slot cur)
    {
        if (status) return cur;
        const bool strided = s_ > 1;
        const int64_t ntrans = o * s_;
        Addr a{};
        a.cnt0 = s_;
        a.cnt1 = 1;
        a.st0 = 1;
        a.st1 = 0;
        a.st2 = n * s_;
        a.es = s_;
        const int64_t limit = dev.max_lds_len[plan.dtype];
        switch (axis_kind(n, s_, o)) {
        case AxisKind::kRows: add_rows(n, ntrans, src, cur, pick_row_variant(n, ntrans)); return cur;
        case AxisKind::kSingle: add_stockham(n, ntrans, a, a, strided, src, cur); return cur;
        case AxisKind::kBluestein: add_bluestein(n, ntrans, a, a, strided, src, cur); return cur;
        case AxisKind::kSplit: break;
        }
        if (s_ == 1) {
            if (const int64_t n1 = six_step_n1(n)) {
                add_six_step(n, n1, o, src, cur);
                return other(cur);
            }
            int64_t n1 = 0, m1 = 0, m2 = 0;
            if (split_six_step_factors(n, o, n1, m1, m2)) {
                add_split_six_step(n, n1, m1, m2, o, src, cur);
                return other(cur);
            }
        }
        // balanced four-step (strided axes, and rows neither six-step form covers):
        // n = n1 * n2, element i = i1*n2 + i2, output k = k1 + n1*k2
        int64_t n1 = 0, n2 = 0;
        double best = 1e300;
        for (int64_t d = 2; d <= limit; ++d) {
            if (n % d) continue;
            const int64_t e = n / d;
            if (e > limit || !is_5smooth(d) || !is_5smooth(e)) continue;
            const double sc = std::fabs(std::log((double)d) - std::log((double)e));
            if (sc < best) {
                best = sc;
                n1 = d;
                n2 = e;
            }
        }
        if (!n1) {
            status = AFFT_ERROR_UNSUPPORTED_SIZE;
            return cur;
        }
        // two-level table for W_n^e: lo[e & (2^sh - 1)], hi[e >> sh]
        int sh = 0;
        while (((int64_t)1 << (2 * sh)) < n) ++sh;
        const int64_t lo_cnt = (int64_t)1 << sh;
        const int64_t hi_cnt = (n + lo_cnt - 1) / lo_cnt;
        void *lo = alloc(lo_cnt * (int64_t)esz());
        void *hi = alloc(hi_cnt * (int64_t)esz());
        if (status) return cur;
        check(gen_power_table(plan.dtype, lo, n, 1, lo_cnt, s));
        check(gen_power_table(plan.dtype, hi, n, lo_cnt, hi_cnt, s));
        const Slot next = other(cur);
        int64_t &room = next == Slot::kWork ? plan.work_elems
                        : next == Slot::kAux2 ? plan.aux2_elems : plan.aux_elems;
        if (next != Slot::kOut) room = std::max(room, o * n * s_);
        // pass A: over i1 (stride n2*s_) for each (c, i2, o), twiddled by W_n^{i2 k1}; k1 takes
        // the position of i1, so the pass may run in place
        Addr ia{};
        ia.cnt0 = s_;
        ia.cnt1 = n2;
        ia.st0 = 1;
        ia.st1 = s_;
        ia.st2 = n * s_;
        ia.es = n2 * s_;
        add_stockham(n1, o * s_ * n2, ia, ia, true, src, cur, n, 1, hi, lo, sh);
        // pass B: over i2 (stride s_) for each (c, k1, o), reading (k1*n2 + i2)*s_ + c and
        // writing X[k1 + n1*k2] to (k1 + n1*k2)*s_ + c
        Addr bi{}, bo{};
        bi.cnt0 = s_;
        bi.cnt1 = n1;
        bi.st0 = 1;
        bi.st1 = n2 * s_;
        bi.st2 = n * s_;
        bi.es = s_;
        bo = bi;
        bo.st1 = s_;
        bo.es = n1 * s_;
        add_stockham(n2, o * s_ * n1, bi, bo, true, cur, next);
        return next;
    }
};

} // namespace

int build_plan(Plan &plan, int dtype, int rank, const int64_t *extents, int axes, int device,
               const DeviceInfo &dev, hipStream_t s)
{
    plan.dtype = dtype;
    plan.rank = rank;
    plan.axes = axes;
    plan.device = device;
    plan.extents.assign(extents, extents + rank);
    plan.total = 1;
    for (int i = 0; i < rank; ++i) plan.total *= extents[i];
    Builder b{plan, dev, s};
    if (plan.total == 0) return AFFT_SUCCESS;
    // Innermost axis first (it reads the input); every later pass works on the output or the
    // work buffer.  A four-step split moves the data between the two, so the first pass writes
    // to whichever of them makes the last pass end in the output: with one split the sequence is
    // rows -> work, pass A in place on work, pass B work -> output, and only one intermediate
    // array is live (a 128 MiB one stays within the 256 MB Infinity Cache).
    auto take = [&](int ax) {
        return (axes == AFFT_AXES_ALL) || ax == rank - 1 || (axes == AFFT_AXES_LAST2 && ax >= rank - 2);
    };
    // the kernels index one transform with 32-bit lengths: a transformed extent of 2^31 or more
    // is refused here, before any length is narrowed
    for (int ax = 0; ax < rank; ++ax)
        if (take(ax) && extents[ax] >= ((int64_t)1 << 31)) return AFFT_ERROR_UNSUPPORTED_SIZE;
    int splits = 0;
    {
        int64_t inner = 1;
        for (int ax = rank - 1; ax >= 0; --ax) {
            if (take(ax) && extents[ax] > 1 &&
                b.axis_kind(extents[ax], inner, plan.total / (extents[ax] * inner)) == Builder::AxisKind::kSplit)
                ++splits;
            inner *= extents[ax];
        }
    }
    Slot src = Slot::kIn;
    Slot cur = splits % 2 ? Slot::kWork : Slot::kOut;
    if (splits) plan.work_elems = plan.total;
    int64_t inner = 1;
    for (int ax = rank - 1; ax >= 0; --ax) {
        const int64_t n = extents[ax];
        if (take(ax) && n > 1) {
            const int64_t o = plan.total / (n * inner);
            cur = b.add_axis(n, inner, o, src, cur);
            src = cur;
            if (b.status) break;
        }
        inner *= n;
    }
    if (b.status == AFFT_SUCCESS && !plan.launches.empty() && cur != Slot::kOut) b.status = AFFT_ERROR_INTERNAL;
    if (b.status == AFFT_SUCCESS && plan.launches.empty()) {
        // every transformed extent is 1: the transform is the identity (Adhoc.hs:45); a copy
        // kernel still applies the scale
        Addr a{};
        a.cnt0 = 1;
        a.cnt1 = 1;
        a.st2 = 1;
        a.es = 1;
        b.add_stockham(1, plan.total, a, a, false, Slot::kIn, Slot::kOut);
    }
    // the work and auxiliary arrays are not allocated here: each execution leases its own
    if (b.status == AFFT_SUCCESS && hipStreamSynchronize(s) != hipSuccess) b.status = AFFT_ERROR_HIP;
    return b.status;
}

static void free_block(void *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

void free_plan(Plan &plan)
{
    for (void *p : plan.allocs) (void)hipFree(p);
    plan.allocs.clear();
    plan.launches.clear();
    plan.tw_cache.clear();
    plan.pow_cache.clear();
    plan.blue_cache.clear();
    if (plan.pool) plan.pool->clear(free_block);
}

static int64_t align256(int64_t b) { return (b + 255) & ~(int64_t)255; }

int64_t workspace_bytes(const Plan &plan)
{
    const int64_t esz = plan.dtype == AFFT_C64 ? 8 : 16;
    return align256(plan.work_elems * esz) + align256(plan.aux_elems * esz) + align256(plan.aux2_elems * esz);
}

Workspace carve_workspace(const Plan &plan, void *base)
{
    const int64_t esz = plan.dtype == AFFT_C64 ? 8 : 16;
    char *p = static_cast<char *>(base);
    Workspace w;
    if (!p) return w;
    if (plan.work_elems) w.work = p;
    p += align256(plan.work_elems * esz);
    if (plan.aux_elems) w.aux = p;
    p += align256(plan.aux_elems * esz);
    if (plan.aux2_elems) w.aux2 = p;
    return w;
}

static void *slot_ptr(Slot s, const void *in, void *out, const Workspace &ws)
{
    switch (s) {
    case Slot::kIn: return const_cast<void *>(in);
    case Slot::kOut: return out;
    case Slot::kWork: return ws.work;
    case Slot::kAux: return ws.aux;
    case Slot::kAux2: return ws.aux2;
    }
    return nullptr;
}

// the launches of a plan with a given workspace
static int enqueue_plan(const Plan &plan, int direction, const void *d_in, void *d_out, double scale,
                        hipStream_t s, const int64_t *in_rows, const int64_t *out_rows,
                        const SignSpec *sg_in, const SignSpec *sg_out, const Workspace &ws);

static FastDiv make_fastdiv(uint64_t d)
{
    int sh = 0;
    while (((uint64_t)1 << sh) < d) ++sh;
    FastDiv f;
    f.m = (((uint64_t)1 << (31 + sh)) / d) + 1;
    f.sh = 31 + sh;
    return f;
}

bool make_sign_spec(int rank, const int64_t *extents, SignSpec &out)
{
    out = SignSpec{};
    if (rank < 1 || rank > 3) return false;
    int64_t total = 1;
    for (int a = 0; a < rank; ++a) total *= extents[a];
    if (total >= ((int64_t)1 << 31)) return false;
    out.on = 1;
    out.rank = rank;
    int64_t stride = total;
    for (int a = 0; a < rank; ++a) {
        out.ext[a] = extents[a];
        stride = extents[a] ? stride / extents[a] : 1;
        out.div[a] = make_fastdiv((uint64_t)std::max<int64_t>(1, stride));
    }
    return true;
}

int exec_plan(const Plan &plan, int direction, const void *d_in, void *d_out, double scale,
              hipStream_t s, const int64_t *in_rows, const int64_t *out_rows, const SignSpec *sg_in,
              const SignSpec *sg_out, const Workspace *ws)
{
    if ((in_rows || out_rows) && (plan.rank != 2 || plan.axes != AFFT_AXES_INNERMOST))
        return AFFT_ERROR_INTERNAL;
    if (plan.total == 0) return AFFT_SUCCESS;
    const int64_t wbytes = workspace_bytes(plan);
    if (ws || wbytes == 0)
        return enqueue_plan(plan, direction, d_in, d_out, scale, s, in_rows, out_rows, sg_in, sg_out,
                            ws ? *ws : Workspace{});
    // lease this execution's own workspace (one per stream in flight, see afft_pool.hpp)
    auto make = [wbytes](void *&mem) {
        mem = nullptr;
        return device_alloc(&mem, (size_t)wbytes) == hipSuccess ? AFFT_SUCCESS : AFFT_ERROR_ALLOC;
    };
    WorkspacePool::Entry *e = nullptr;
    const bool capturing = stream_capturing(s);
    int st = capturing ? plan.pool->acquire_pinned(make, e) : plan.pool->acquire(s, make, e);
    if (st) return st;
    st = enqueue_plan(plan, direction, d_in, d_out, scale, s, in_rows, out_rows, sg_in, sg_out,
                      carve_workspace(plan, e->res));
    plan.pool->release(e, s, true);
    return st;
}

static int enqueue_plan(const Plan &plan, int direction, const void *d_in, void *d_out, double scale,
                        hipStream_t s, const int64_t *in_rows, const int64_t *out_rows,
                        const SignSpec *sg_in, const SignSpec *sg_out, const Workspace &ws)
{
    for (const Launch &l : plan.launches) {
        const Slot used[2] = {l.src, l.dst};
        for (Slot u : used)
            if ((u == Slot::kWork && !ws.work) || (u == Slot::kAux && !ws.aux) || (u == Slot::kAux2 && !ws.aux2))
                return AFFT_ERROR_INTERNAL;
    }
    // in place (d_in == d_out) needs no private copy: every pass that reads the input either runs
    // in place, each workgroup reading all its elements before writing them, or writes the work
    // buffer, and the permuting passes read only the plan's own buffers
    const void *src_in = d_in;
    const int conj = direction > 0 ? 1 : 0;
    const size_t n = plan.launches.size();
    // streaming accesses for the plan's input (read once) and its final output (written once);
    // intermediates keep the default policy so the next pass finds them in L2 / Infinity Cache
    static const bool nt_enabled = [] {
        const char *v = std::getenv("AFFT_NT");
        return !(v && v[0] == '0');
    }();
    int st = AFFT_SUCCESS;
    for (size_t i = 0; i < n && st == AFFT_SUCCESS; ++i) {
        const Launch &l = plan.launches[i];
        const bool first = i == 0, last = i + 1 == n;
        hipError_t e;
        if (l.kind == Launch::kStockham || l.kind == Launch::kRegCol) {
            FftArgs a = l.fa;
            a.in = slot_ptr(l.src, src_in, d_out, ws);
            a.out = slot_ptr(l.dst, src_in, d_out, ws);
            a.nt_in = nt_enabled && l.src == Slot::kIn;
            a.nt_out = nt_enabled && last && l.dst == Slot::kOut;
            a.conj_in = first ? conj : 0;
            a.conj_out = last ? conj : 0;
            a.scale = last ? scale : 1.0;
            if (l.src == Slot::kIn) a.ain.toff = in_rows;
            if (l.dst == Slot::kOut) a.aout.toff = out_rows;
            if (first && sg_in) a.sg_in = *sg_in;
            if (last && sg_out) a.sg_out = *sg_out;
            e = l.kind == Launch::kRegCol ? launch_regcol(plan.dtype, a, l.grid, s)
                                          : launch_stockham(plan.dtype, a, l.block, l.grid, l.lds, s);
        } else if (l.kind == Launch::kRows) {
            RowArgs a = l.ra;
            a.in = slot_ptr(l.src, src_in, d_out, ws);
            a.out = slot_ptr(l.dst, src_in, d_out, ws);
            a.nt_in = nt_enabled && l.src == Slot::kIn;
            a.nt_out = nt_enabled && last && l.dst == Slot::kOut;
            a.conj_in = first ? conj : 0;
            a.conj_out = last ? conj : 0;
            a.scale = last ? scale : 1.0;
            if (l.row_variant == kRowTransposed) {
                // reads the plan's own buffer; the transposed store is per transform
                if (l.src == Slot::kIn) return AFFT_ERROR_INTERNAL;
                if (l.dst == Slot::kOut) a.tr_toff = out_rows;
            } else if ((in_rows || out_rows) && (l.src == Slot::kIn || l.dst == Slot::kOut)) {
                // one offset table serves both sides: rows at the same offsets in and out
                if (in_rows != out_rows || l.src != Slot::kIn || l.dst != Slot::kOut) return AFFT_ERROR_INTERNAL;
                a.roffs = in_rows;
            }
            if (first && sg_in) a.sg_in = *sg_in;
            if (last && sg_out) a.sg_out = *sg_out;
            e = launch_rows(plan.dtype, l.row_n, l.row_variant, a, l.grid, s);
        } else if (l.kind == Launch::kTranspose) {
            TransposeArgs a = l.ta;
            a.in = slot_ptr(l.src, src_in, d_out, ws);
            a.out = slot_ptr(l.dst, src_in, d_out, ws);
            a.nt_out = nt_enabled && last && l.dst == Slot::kOut;
            a.conj_out = last ? conj : 0;
            a.scale = last ? scale : 1.0;
            if (l.src == Slot::kIn || first) return AFFT_ERROR_INTERNAL;
            if (l.dst == Slot::kOut) a.toff = out_rows;
            if (last && sg_out) a.sg_out = *sg_out;
            e = launch_transpose(plan.dtype, a, s);
        } else if (l.kind == Launch::kChirp) {
            ChirpArgs a = l.ca;
            a.in = slot_ptr(l.src, src_in, d_out, ws);
            a.out = slot_ptr(l.dst, src_in, d_out, ws);
            a.nt_in = nt_enabled && l.src == Slot::kIn;
            a.nt_out = nt_enabled && last && l.dst == Slot::kOut;
            a.conj_in = first ? conj : 0;
            a.conj_out = last ? conj : 0;
            a.scale = last ? scale : 1.0;
            if (l.src == Slot::kIn) a.a.toff = in_rows;
            if (l.dst == Slot::kOut) a.a.toff = out_rows;
            if (first && sg_in) a.sg = *sg_in;
            if (last && sg_out) a.sg = *sg_out;
            e = launch_chirp(plan.dtype, a, s);
        } else {
            BluesteinArgs a = l.ba;
            a.in = slot_ptr(l.src, src_in, d_out, ws);
            a.out = slot_ptr(l.dst, src_in, d_out, ws);
            a.nt_in = nt_enabled && l.src == Slot::kIn;
            a.nt_out = nt_enabled && last && l.dst == Slot::kOut;
            a.conj_in = first ? conj : 0;
            a.conj_out = last ? conj : 0;
            a.scale = last ? scale : 1.0;
            if (l.src == Slot::kIn) a.ain.toff = in_rows;
            if (l.dst == Slot::kOut) a.aout.toff = out_rows;
            if (first && sg_in) a.sg_in = *sg_in;
            if (last && sg_out) a.sg_out = *sg_out;
            e = launch_bluestein(plan.dtype, a, l.block, l.grid, l.lds, s);
        }
        if (e != hipSuccess) st = AFFT_ERROR_HIP;
    }
    return st;
}

// ----------------------------------------------------------------------------------------------
// Device info and the plan cache.

static std::mutex g_dev_mu;
static std::map<int, DeviceInfo> g_dev;

int device_info(int device, DeviceInfo &out)
{
    std::lock_guard<std::mutex> lk(g_dev_mu);
    auto it = g_dev.find(device);
    if (it != g_dev.end()) {
        out = it->second;
        return AFFT_SUCCESS;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return AFFT_ERROR_NO_DEVICE;
    if (device < 0 || device >= count) return AFFT_ERROR_INVALID_VALUE;
    DeviceInfo d{};
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
        return AFFT_ERROR_HIP;
    d.lds_bytes = lds > 0 ? (size_t)lds : 65536;
    int cus = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    d.cus = cus;
    // leave room for the padding of the largest power of two that fits
    d.max_lds_len[AFFT_C64] = lds_limit_elems(AFFT_C64, d.lds_bytes);
    d.max_lds_len[AFFT_C128] = lds_limit_elems(AFFT_C128, d.lds_bytes);
    g_dev[device] = d;
    out = d;
    return AFFT_SUCCESS;
}

struct DeviceGuard {
    int old = -1;
    bool ok = false;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&old) != hipSuccess) return;
        ok = old == dev || hipSetDevice(dev) == hipSuccess;
    }
    ~DeviceGuard()
    {
        if (ok && old >= 0) (void)hipSetDevice(old);
    }
};

int create_plan(Plan *&out, int dtype, int rank, const int64_t *extents, int axes, int device)
{
    reset_alloc_count();
    DeviceInfo dev;
    int st = device_info(device, dev);
    if (st) return st;
    DeviceGuard g(device);
    if (!g.ok) return AFFT_ERROR_HIP;
    hipStream_t s;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return AFFT_ERROR_HIP;
    Plan *p = new (std::nothrow) Plan();
    if (!p) {
        (void)hipStreamDestroy(s);
        return AFFT_ERROR_ALLOC;
    }
    st = build_plan(*p, dtype, rank, extents, axes, device, dev, s);
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    if (st) {
        free_plan(*p);
        delete p;
        return st;
    }
    out = p;
    return AFFT_SUCCESS;
}

using Key = std::tuple<int, int, int, int, std::vector<int64_t>>;
static std::mutex g_cache_mu;
static std::map<Key, Plan *> g_cache;

int cached_plan(const Plan *&out, int dtype, int rank, const int64_t *extents, int axes, int device)
{
    Key k{device, dtype, rank, axes, std::vector<int64_t>(extents, extents + rank)};
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto it = g_cache.find(k);
    if (it != g_cache.end()) {
        out = it->second;
        return AFFT_SUCCESS;
    }
    Plan *p = nullptr;
    int st = create_plan(p, dtype, rank, extents, axes, device);
    if (st) return st;
    g_cache[k] = p;
    out = p;
    return AFFT_SUCCESS;
}

void clear_cache()
{
    std::lock_guard<std::mutex> lk(g_cache_mu);
    for (auto &kv : g_cache) {
        int old = -1;
        (void)hipGetDevice(&old);
        (void)hipSetDevice(kv.second->device);
        (void)hipDeviceSynchronize();
        free_plan(*kv.second);
        delete kv.second;
        if (old >= 0) (void)hipSetDevice(old);
    }
    g_cache.clear();
}

int run_plan(const Plan &p, int direction, const void *in, void *out, double scale, void *stream)
{
    DeviceGuard g(p.device);
    if (!g.ok) return AFFT_ERROR_HIP;
    return exec_plan(p, direction, in, out, scale, (hipStream_t)stream);
}

int current_device(int &dev)
{
    if (hipGetDevice(&dev) != hipSuccess) return AFFT_ERROR_NO_DEVICE;
    return AFFT_SUCCESS;
}

// Table helpers for the engines outside the axis planner (ragged batches, spectral
// convolution): the allocations belong to `owner` and are released by free_plan(owner).
void *plan_alloc(Plan &owner, const DeviceInfo &dev, int64_t bytes, int &status)
{
    Builder b{owner, dev, nullptr};
    void *p = b.alloc(bytes);
    if (b.status) status = b.status;
    return p;
}

int table_pass_twiddles(Plan &owner, const DeviceInfo &dev, hipStream_t s, int64_t n,
                        const LdsLayout &L, PassTable &pt, const void *&tw)
{
    Builder b{owner, dev, s};
    tw = b.twiddles(n, L, pt);
    return b.status;
}

int table_power(Plan &owner, const DeviceInfo &dev, hipStream_t s, int64_t L, int64_t step,
                int64_t count, const void *&out)
{
    Builder b{owner, dev, s};
    void *buf = b.alloc(count * (int64_t)b.esz());
    if (!b.status) b.check(gen_power_table(owner.dtype, buf, L, step, count, s));
    out = buf;
    return b.status;
}

int table_bluestein(Plan &owner, const DeviceInfo &dev, hipStream_t s, int64_t n, int64_t M,
                    const void *&chirp, const void *&filt)
{
    Builder b{owner, dev, s};
    b.bluestein_tables(n, M, chirp, filt);
    return b.status;
}

} // namespace afft

// ==============================================================================================
// C ABI

using namespace afft;

static bool valid_dtype(int d) { return d == AFFT_C64 || d == AFFT_C128; }
static bool valid_mode(int m) { return m == AFFT_FORWARD || m == AFFT_REVERSE || m == AFFT_INVERSE; }

extern "C" const char *afft_strerror(int status)
{
    switch (status) {
    case AFFT_SUCCESS: return "success";
    case AFFT_ERROR_INVALID_VALUE: return "invalid argument";
    case AFFT_ERROR_UNSUPPORTED_RANK: return "unsupported rank";
    case AFFT_ERROR_UNSUPPORTED_DTYPE: return "unsupported element type";
    case AFFT_ERROR_UNSUPPORTED_SIZE: return "unsupported transform length";
    case AFFT_ERROR_ALLOC: return "allocation failed";
    case AFFT_ERROR_HIP: return "HIP runtime error";
    case AFFT_ERROR_NO_DEVICE: return "no usable device";
    case AFFT_ERROR_IO: return "plan-cache file error";
    case AFFT_ERROR_INTERNAL: return "internal error";
    case AFFT_ERROR_SCRATCH_TOO_SMALL: return "the plan's intermediate arrays exceed the allocation-free scratch";
    default: return "unknown status";
    }
}

extern "C" int afft_version(void) { return 10000; }

static int check_shape(int dtype, int rank, const int64_t *extents, int max_rank)
{
    if (!valid_dtype(dtype)) return AFFT_ERROR_UNSUPPORTED_DTYPE;
    if (rank < 1 || rank > max_rank) return AFFT_ERROR_UNSUPPORTED_RANK;
    if (!extents) return AFFT_ERROR_INVALID_VALUE;
    for (int i = 0; i < rank; ++i)
        if (extents[i] < 0) return AFFT_ERROR_INVALID_VALUE;
    // the array's size in bytes must fit an int64_t (an empty array always does)
    int64_t bytes = 16;
    for (int i = 0; i < rank; ++i) {
        if (extents[i] == 0) return AFFT_SUCCESS;
        if (__builtin_mul_overflow(bytes, extents[i], &bytes)) return AFFT_ERROR_UNSUPPORTED_SIZE;
    }
    return AFFT_SUCCESS;
}

static int64_t total_of(int rank, const int64_t *e)
{
    int64_t t = 1;
    for (int i = 0; i < rank; ++i) t *= e[i];
    return t;
}

// The front-end operation: mode semantics of FFT.hs.  Inverse divides by the product of the
// extents from axis `div_from` inward: the innermost extent for fft (FFT.hs:69), every extent for
// fft1D/fft2D/fft3D (FFT.hs:98,125,155), the two innermost for the batched fft2D.
static int front_end(int mode, int dtype, int rank, const int64_t *extents, int axes, int div_from,
                     const void *in, void *out, void *stream)
{
    if (!valid_mode(mode)) return AFFT_ERROR_INVALID_VALUE;
    int st = check_shape(dtype, rank, extents, 8);
    if (st) return st;
    if (total_of(rank, extents) == 0) return AFFT_SUCCESS;
    if (!in || !out) return AFFT_ERROR_INVALID_VALUE;
    int dev;
    if ((st = current_device(dev))) return st;
    const Plan *p = nullptr;
    if ((st = cached_plan(p, dtype, rank, extents, axes, dev))) return st;
    const int dir = mode == AFFT_FORWARD ? -1 : 1;
    const int64_t divisor = total_of(rank - div_from, extents + div_from);
    const double scale = (mode == AFFT_INVERSE && divisor > 0) ? 1.0 / (double)divisor : 1.0;
    return run_plan(*p, dir, in, out, scale, stream);
}

extern "C" int afft_fft(int mode, int dtype, int rank, const int64_t *extents, const void *d_in,
                        void *d_out, void *stream)
{
    int st = check_shape(dtype, rank, extents, 8);
    if (st) return st;
    return front_end(mode, dtype, rank, extents, AFFT_AXES_INNERMOST, rank - 1, d_in, d_out, stream);
}

extern "C" int afft_fft1D(int mode, int dtype, int64_t n, const void *d_in, void *d_out, void *stream)
{
    const int64_t e[1] = {n};
    return front_end(mode, dtype, 1, e, AFFT_AXES_ALL, 0, d_in, d_out, stream);
}

extern "C" int afft_fft2D(int mode, int dtype, int64_t h, int64_t w, const void *d_in, void *d_out,
                          void *stream)
{
    const int64_t e[2] = {h, w};
    return front_end(mode, dtype, 2, e, AFFT_AXES_ALL, 0, d_in, d_out, stream);
}

extern "C" int afft_fft3D(int mode, int dtype, int64_t d, int64_t h, int64_t w, const void *d_in,
                          void *d_out, void *stream)
{
    const int64_t e[3] = {d, h, w};
    return front_end(mode, dtype, 3, e, AFFT_AXES_ALL, 0, d_in, d_out, stream);
}

extern "C" int afft_fft2D_many(int mode, int dtype, int64_t batch, int64_t h, int64_t w,
                               const void *d_in, void *d_out, void *stream)
{
    const int64_t e[3] = {batch, h, w};
    return front_end(mode, dtype, 3, e, AFFT_AXES_LAST2, 1, d_in, d_out, stream);
}

extern "C" int afft_plan_create(afft_plan **plan, int dtype, int rank, const int64_t *extents, int axes,
                                int device)
{
    if (!plan) return AFFT_ERROR_INVALID_VALUE;
    *plan = nullptr;
    int st = check_shape(dtype, rank, extents, 8);
    if (st) return st;
    if (axes != AFFT_AXES_INNERMOST && axes != AFFT_AXES_ALL && axes != AFFT_AXES_LAST2)
        return AFFT_ERROR_INVALID_VALUE;
    Plan *p = nullptr;
    st = create_plan(p, dtype, rank, extents, axes, device);
    if (st) return st;
    *plan = reinterpret_cast<afft_plan *>(p);
    return AFFT_SUCCESS;
}

extern "C" int afft_plan_cached(const afft_plan **plan, int dtype, int rank, const int64_t *extents,
                                int axes, int device)
{
    if (!plan) return AFFT_ERROR_INVALID_VALUE;
    *plan = nullptr;
    int st = check_shape(dtype, rank, extents, 8);
    if (st) return st;
    if (axes != AFFT_AXES_INNERMOST && axes != AFFT_AXES_ALL && axes != AFFT_AXES_LAST2)
        return AFFT_ERROR_INVALID_VALUE;
    const Plan *p = nullptr;
    st = cached_plan(p, dtype, rank, extents, axes, device);
    if (st) return st;
    *plan = reinterpret_cast<const afft_plan *>(p);
    return AFFT_SUCCESS;
}

extern "C" int afft_exec(const afft_plan *plan, int direction, const void *d_in, void *d_out,
                         double scale, void *stream)
{
    if (!plan || (direction != -1 && direction != 1)) return AFFT_ERROR_INVALID_VALUE;
    const Plan *p = reinterpret_cast<const Plan *>(plan);
    if (p->total == 0) return AFFT_SUCCESS;
    if (!d_in || !d_out) return AFFT_ERROR_INVALID_VALUE;
    return run_plan(*p, direction, d_in, d_out, scale, stream);
}

extern "C" int afft_plan_destroy(afft_plan *plan)
{
    if (!plan) return AFFT_ERROR_INVALID_VALUE;
    Plan *p = reinterpret_cast<Plan *>(plan);
    {
        DeviceGuard g(p->device);
        (void)hipDeviceSynchronize();
        free_plan(*p);
    }
    delete p;
    return AFFT_SUCCESS;
}

extern "C" int afft_cache_clear(void)
{
    clear_cache();
    return AFFT_SUCCESS;
}

extern "C" int afft_plan_info(const afft_plan *plan, int *launches, int64_t *device_bytes)
{
    if (!plan) return AFFT_ERROR_INVALID_VALUE;
    const Plan *p = reinterpret_cast<const Plan *>(plan);
    if (launches) *launches = (int)p->launches.size();
    if (device_bytes) *device_bytes = p->device_bytes + workspace_bytes(*p);
    return AFFT_SUCCESS;
}

extern "C" int afft_describe(int64_t n, int dtype, int *radices, int max_passes, int *n_passes,
                             int *elems_per_thread, int64_t *bluestein_len)
{
    if (!valid_dtype(dtype)) return AFFT_ERROR_UNSUPPORTED_DTYPE;
    if (n < 1) return AFFT_ERROR_INVALID_VALUE;
    // the single-workgroup limit on a 160 KiB LDS (gfx950)
    const int64_t limit = lds_limit_elems(dtype, 160 * 1024);
    int64_t len = n, blue = 0;
    if (!is_5smooth(n)) {
        blue = bluestein_length(n);
        len = blue;
    }
    if (len > limit) return AFFT_ERROR_UNSUPPORTED_SIZE;
    LdsLayout L{};
    if (!choose_layout(len, dtype, L)) return AFFT_ERROR_UNSUPPORTED_SIZE;
    if (n_passes) *n_passes = L.npass;
    if (elems_per_thread) *elems_per_thread = L.E;
    if (bluestein_len) *bluestein_len = blue;
    if (radices)
        for (int i = 0; i < L.npass && i < max_passes; ++i) radices[i] = L.radix[i];
    return AFFT_SUCCESS;
}
