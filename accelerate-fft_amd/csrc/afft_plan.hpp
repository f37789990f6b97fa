// afft_plan.hpp -- the plan representation shared by the planner and the extension entry points.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <map>
#include <memory>
#include <utility>
#include <vector>

#include "afft_internal.hpp"
#include "afft_pool.hpp"

namespace afft {

// Layout of one single-workgroup transform: E elements per thread and the pass radices.
struct LdsLayout {
    int E;
    int npass;
    int radix[kMaxPasses];
};

struct DeviceInfo {
    size_t lds_bytes;        // LDS a workgroup may allocate
    int cus;                 // compute units
    int64_t max_lds_len[2];  // longest single-workgroup transform per dtype
};

enum class Slot { kIn, kOut, kWork, kAux, kAux2 };

struct Launch {
    enum Kind { kStockham, kBluestein, kRows, kChirp, kRegCol, kTranspose } kind;
    Slot src, dst;
    FftArgs fa;
    BluesteinArgs ba;
    RowArgs ra;
    ChirpArgs ca;
    TransposeArgs ta;
    int row_n = 0, row_variant = 0;
    int block;
    int64_t grid;
    size_t lds;
};

// The device arrays one execution of a plan writes besides its output: the work array of split
// axes and the two auxiliary arrays of multi-kernel Bluestein transforms.  They belong to the
// execution, not to the plan: each execution leases its own (afft_pool.hpp), or the caller brings
// them (ragged batches, the allocation-free entry point).
struct Workspace {
    void *work = nullptr;
    void *aux = nullptr;
    void *aux2 = nullptr;
};
using WorkspacePool = ExecPool<void *>; // a lease is one device block, carved into a Workspace

// Plan::flags: the plan runs on gathered rows that start at arbitrary offsets (ragged batches),
// so short row segments of a transposed store would straddle cache lines
constexpr int kPlanUnalignedRows = 1;

struct Plan {
    int dtype = 0, rank = 0, axes = 0, device = 0;
    int flags = 0;
    std::vector<int64_t> extents;
    int64_t total = 0;
    std::vector<Launch> launches;
    std::vector<void *> allocs;
    int64_t device_bytes = 0;
    std::map<int64_t, const void *> tw_cache;
    std::map<int64_t, const void *> pow_cache;
    std::map<int64_t, std::pair<const void *, const void *>> blue_cache;
    // elements of each per-execution array (0: the plan needs none of it)
    int64_t work_elems = 0;
    int64_t aux_elems = 0;
    int64_t aux2_elems = 0;  // partner of aux for four-step splits inside a Bluestein transform
    std::shared_ptr<WorkspacePool> pool = std::make_shared<WorkspacePool>();
};

// Bytes of one execution's workspace (the three arrays, each 256-byte aligned), 0 if none.
int64_t workspace_bytes(const Plan &plan);
// The three arrays laid out in a block of workspace_bytes(plan) at base.
Workspace carve_workspace(const Plan &plan, void *base);

// Single-workgroup layout of length n: fewest passes, then E closest to 16; optional limits on
// elements per thread and threads per transform.
bool choose_layout(int64_t n, int dtype, LdsLayout &out, int maxE = 1 << 30, int maxT = -1);
int64_t bluestein_length(int64_t n);
int64_t lds_limit_elems(int dtype, size_t lds_bytes);
int device_info(int device, DeviceInfo &out);
int build_plan(Plan &plan, int dtype, int rank, const int64_t *extents, int axes, int device,
               const DeviceInfo &dev, hipStream_t s);
void free_plan(Plan &plan);
// in_rows / out_rows (optional, plans of shape (count, L) along the innermost axis only): element
// offsets of the count rows in d_in / d_out, for the rows of one length of a ragged batch.
// sg_in / sg_out (optional): a centring sign applied on the first pass's loads / the last pass's
// stores (fused DFT/Centre.hs operations).
// ws (optional): the workspace to use; without it the execution leases one from the plan's pool.
int exec_plan(const Plan &plan, int direction, const void *d_in, void *d_out, double scale,
              hipStream_t s, const int64_t *in_rows = nullptr, const int64_t *out_rows = nullptr,
              const SignSpec *sg_in = nullptr, const SignSpec *sg_out = nullptr,
              const Workspace *ws = nullptr);
// the sign (-1)^(sum of the multi-index) over a row-major array of rank <= 3 (false: too large)
bool make_sign_spec(int rank, const int64_t *extents, SignSpec &out);
int create_plan(Plan *&out, int dtype, int rank, const int64_t *extents, int axes, int device);
int cached_plan(const Plan *&out, int dtype, int rank, const int64_t *extents, int axes, int device);
int run_plan(const Plan &p, int direction, const void *in, void *out, double scale, void *stream);
int current_device(int &dev);
// hipMalloc through the library's failure-injection point (AFFT_TEST_FAIL_ALLOC), and the reset
// of its per-thread count at the start of a plan creation
hipError_t device_alloc(void **p, size_t bytes);
void reset_alloc_count();
void clear_cache();
void *plan_alloc(Plan &owner, const DeviceInfo &dev, int64_t bytes, int &status);
int table_pass_twiddles(Plan &owner, const DeviceInfo &dev, hipStream_t s, int64_t n,
                        const LdsLayout &L, PassTable &pt, const void *&tw);
int table_power(Plan &owner, const DeviceInfo &dev, hipStream_t s, int64_t L, int64_t step,
                int64_t count, const void *&out);
int table_bluestein(Plan &owner, const DeviceInfo &dev, hipStream_t s, int64_t n, int64_t M,
                    const void *&chirp, const void *&filt);
bool is_5smooth(int64_t n);
// The autotuner (afft_tune.cpp).  The planner asks tune_choice for every layout decision it
// makes for a length, with the candidates it accepts (the first is its default); the answer is a
// choice forced by the autotuner on this thread, else the tuned one from the table, else -1.
enum TuneKind { kTuneRows = 0, kTuneStrided = 1, kTuneRegcol = 2, kTuneRagged = 3, kTuneRaggedSplit = 4, kTuneKinds = 5 };
struct TunePoint {
    int kind, dtype;
    int64_t n;
    std::vector<int> cands;
};
int tune_choice(int kind, int dtype, int64_t n, int device, const std::vector<int> &cands);
// Record the decision points the planner meets on this thread into `points` (null: stop).
void tune_record(std::vector<TunePoint> *points);
std::vector<TunePoint> *tune_recording();
// Time each candidate of each point in turn, the choices made so far forced, and keep the fastest
// in the table; time_with_forced builds and times the work, returning microseconds (< 0: failed).
int tune_points(const std::vector<TunePoint> &points, int device, const std::function<double()> &time_with_forced);
// Device time of one call (microseconds, < 0: failed): the median of three batches of five.
double tune_time_calls(const std::function<int(hipStream_t)> &call, hipStream_t s);

} // namespace afft
