// afft_pool.hpp -- per-execution device resources of a plan, leased per stream.
//
// A plan is shared by every thread and stream that uses it (the plan cache hands one plan to all
// callers of a key), but the intermediate arrays of one execution -- the work array of a split
// axis, the auxiliary arrays of a multi-kernel Bluestein transform, the side stream of a ragged
// batch -- must not be shared by two executions in flight at once.  An ExecPool keeps a few such
// resource sets per plan and leases one to each execution:
//
//   * an idle set last used on the caller's stream is taken first: the stream itself orders the
//     two executions, so nothing is recorded or waited on (a stream that keeps one plan busy pays
//     no synchronisation at all);
//   * else an unused set, or one whose recorded last execution has finished on the device;
//   * else a new set, up to kMaxLeases;
//   * else an idle set still in flight on another stream: an event is recorded on that stream
//     (after the executions that used the set) and the caller's stream waits on it (correct,
//     only less concurrent);
//   * else (every set is being enqueued by another host thread) the caller waits for one.
//
// So executions on different streams never share a set while either is in flight, and a single
// stream keeps reusing one set.  This settles the open question of the cuFFT plugin,
// whose handle carries the stream and the work area together (LLVM/PTX.hs:121-124,
// PTX/Plans.hs:58-63).
//
// A stream under graph capture cannot wait on or record events from outside the capture, and
// the captured graph replays later at times the library does not see.  Such an execution gets a
// pinned set of its own, which is never leased again and lives as long as the plan.
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <functional>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/afft.h"

namespace afft {

constexpr int kMaxLeases = 4;

// true when s is recording a graph capture (its work is replayed later, not run now)
inline bool stream_capturing(hipStream_t s)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess) return false;
    return cs == hipStreamCaptureStatusActive;
}


template <class R> class ExecPool {
public:
    struct Entry {
        R res{};
        hipEvent_t done = nullptr;    // recorded after the executions that used res (see dirty)
        hipStream_t last = nullptr;   // the stream of the last execution that used res
        bool used = false;            // some execution has used res
        bool dirty = false;           // executions on `last` since `done` was recorded
        bool busy = false;            // leased, its execution not yet enqueued
        bool pinned = false;          // owned by a captured graph or retired after an error
    };
    using Make = std::function<int(R &)>;
    using Free = std::function<void(R &)>;

    // Lease a set for one execution on stream s (not under capture).  `make` creates the
    // resources of a new set and returns a status.
    int acquire(hipStream_t s, const Make &make, Entry *&out)
    {
        out = nullptr;
        const bool per_thread = s == hipStreamPerThread;
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            Entry *pick = nullptr;
            bool wait = false;
            int unpinned = 0;
            for (auto &e : entries_)
                if (!e->pinned) ++unpinned;
            // 1. an idle set whose executions all went to this stream: the stream orders them
            if (!per_thread)
                for (auto &e : entries_)
                    if (!e->pinned && !e->busy && e->used && e->last == s) {
                        pick = e.get();
                        break;
                    }
            // 2. a set never used, or one whose recorded last execution has finished
            if (!pick)
                for (auto &e : entries_)
                    if (!e->pinned && !e->busy &&
                        (!e->used || (!e->dirty && hipEventQuery(e->done) == hipSuccess))) {
                        pick = e.get();
                        break;
                    }
            // 3. a new set
            if (!pick && unpinned < kMaxLeases) {
                std::unique_ptr<Entry> e(new (std::nothrow) Entry());
                if (!e) return AFFT_ERROR_ALLOC;
                if (hipEventCreateWithFlags(&e->done, hipEventDisableTiming) != hipSuccess) return AFFT_ERROR_HIP;
                const int st = make(e->res);
                if (st) {
                    (void)hipEventDestroy(e->done);
                    return st;
                }
                pick = e.get();
                entries_.push_back(std::move(e));
            }
            // 4. an idle set in flight on another stream: this stream waits for it on the device.
            // Executions of a set on one stream are not recorded one by one; the event is recorded
            // on that stream now, after them (never into a capture in progress there).
            if (!pick)
                for (auto &e : entries_) {
                    if (e->pinned || e->busy) continue;
                    if (e->dirty) {
                        if (stream_capturing(e->last)) continue;
                        if (hipEventRecord(e->done, e->last) != hipSuccess) continue;
                        e->dirty = false;
                    }
                    pick = e.get();
                    wait = true;
                    break;
                }
            if (!pick) {
                // every idle set sits on a stream that cannot take an event now (a capture in
                // progress, or a stream that is gone): with none busy, nobody would wake this
                // thread, so take one more set beyond the bound
                bool any_busy = false;
                for (auto &e : entries_)
                    if (!e->pinned && e->busy) any_busy = true;
                if (!any_busy) {
                    std::unique_ptr<Entry> e(new (std::nothrow) Entry());
                    if (!e) return AFFT_ERROR_ALLOC;
                    if (hipEventCreateWithFlags(&e->done, hipEventDisableTiming) != hipSuccess) return AFFT_ERROR_HIP;
                    const int st = make(e->res);
                    if (st) {
                        (void)hipEventDestroy(e->done);
                        return st;
                    }
                    pick = e.get();
                    entries_.push_back(std::move(e));
                }
            }
            if (pick) {
                pick->busy = true;
                lk.unlock();
                if (wait && hipStreamWaitEvent(s, pick->done, 0) != hipSuccess) {
                    release(pick, nullptr, false);
                    return AFFT_ERROR_HIP;
                }
                out = pick;
                return AFFT_SUCCESS;
            }
            cv_.wait(lk);
        }
    }

    // A set of its own for an execution captured into a graph: never leased again.  `make` runs
    // with the capture mode of this thread relaxed, so it may allocate during a capture.
    int acquire_pinned(const Make &make, Entry *&out)
    {
        out = nullptr;
        std::unique_ptr<Entry> e(new (std::nothrow) Entry());
        if (!e) return AFFT_ERROR_ALLOC;
        hipStreamCaptureMode mode = hipStreamCaptureModeRelaxed;
        (void)hipThreadExchangeStreamCaptureMode(&mode);
        const int st = make(e->res);
        (void)hipThreadExchangeStreamCaptureMode(&mode);
        if (st) return st;
        e->pinned = true;
        e->busy = true;
        std::lock_guard<std::mutex> lk(mu_);
        out = e.get();
        entries_.push_back(std::move(e));
        return AFFT_SUCCESS;
    }

    // Return a lease after its execution was enqueued on s.  No event is recorded here: a set
    // that stays on one stream never needs one, and another stream that wants the set records
    // it on `last` then.  The per-thread default stream names a different stream in every
    // thread, so an execution there records its event at once.  Pinned sets stay out of
    // circulation.
    void release(Entry *e, hipStream_t s, bool enqueued)
    {
        if (!e) return;
        if (!e->pinned && enqueued) {
            e->used = true;
            e->last = s;
            if (s == hipStreamPerThread) {
                if (hipEventRecord(e->done, s) == hipSuccess) e->dirty = false;
                else e->pinned = true; // its last use cannot be tracked: retire it
            } else {
                e->dirty = true;
            }
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            e->busy = false;
        }
        cv_.notify_one();
    }

    // Free one pinned set (its graph is gone and has finished on the device).
    void drop(Entry *e, const Free &fr)
    {
        std::lock_guard<std::mutex> lk(mu_);
        for (size_t i = 0; i < entries_.size(); ++i)
            if (entries_[i].get() == e) {
                fr(e->res);
                if (e->done) (void)hipEventDestroy(e->done);
                entries_.erase(entries_.begin() + (ptrdiff_t)i);
                return;
            }
    }

    // Free every set; the caller has synchronised the device.
    void clear(const Free &fr)
    {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto &e : entries_) {
            fr(e->res);
            if (e->done) (void)hipEventDestroy(e->done);
        }
        entries_.clear();
    }

    size_t size()
    {
        std::lock_guard<std::mutex> lk(mu_);
        return entries_.size();
    }

private:
    std::mutex mu_;
    std::condition_variable cv_;
    std::vector<std::unique_ptr<Entry>> entries_;
};

} // namespace afft
