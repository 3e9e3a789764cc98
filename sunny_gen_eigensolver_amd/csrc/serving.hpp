// Serving session: a request loop over resident solvers. Requests (a batch of
// pencils of one supported order, e.g. the q-points of one query) are queued
// as they arrive. A scheduler thread cuts each request into slices of at most
// `slice` pencils (0 = whole requests) and puts one slice on every free slot
// (a HIP stream with its own solver handles and workspaces), choosing the
// request by policy: arrival order, or the least remaining work first
// (remaining pencils x n^3). A request staged through host buffers takes its
// device footprint from the session's HBM arena when its first slice starts
// and gives it back when its last slice is done. Slices run concurrently on
// their slots' streams: inputs in, solve, outputs out, then an event marks
// the slice done. Every slice waits for an event the caller's stream
// recorded at submission, so buffers the caller produced asynchronously are
// ready before the session reads or writes them.
//
// No counterpart in the reference (it launches one batch per call,
// src/unified/unified_solver.cpp:1063-1125).
#pragma once

#include "solver.hpp"

#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace sgeig {

// First-fit allocator over one device allocation (offsets, 256-byte aligned,
// free blocks coalesced). Not thread-safe: the session serialises it.
class Arena {
public:
    explicit Arena(size_t bytes);
    // Offset of a block of `bytes`, or SIZE_MAX when nothing fits.
    size_t alloc(size_t bytes);
    void release(size_t offset);
    size_t in_use() const { return in_use_; }
    size_t size() const { return size_; }

private:
    size_t size_ = 0;
    size_t in_use_ = 0;
    std::map<size_t, size_t> free_;   // offset -> size
    std::map<size_t, size_t> used_;   // offset -> size
};

struct ServingStats {
    int64_t submitted = 0;
    int64_t completed = 0;
    size_t hbm_peak = 0;          // arena bytes in use, high-water mark
    int max_concurrent = 0;       // slices running at once, high-water
};

enum class Policy { Fifo = 0, ShortestRemaining = 1 };

class ServingSession {
public:
    ServingSession(int device, Mode mode, const std::vector<int>& orders, int max_count, int slots,
                   size_t hbm_budget, CacheMode cache);
    ~ServingSession();
    ServingSession(const ServingSession&) = delete;
    ServingSession& operator=(const ServingSession&) = delete;

    // Queue a request of `count` pencils of order n. With on_device the five
    // pointers are caller-owned device buffers that stay valid until the
    // request completes; otherwise they are host buffers (pinned memory gives
    // asynchronous copies) and the session stages them through its arena.
    // The session's work on the request waits for what `ready` (the stream
    // that produced the buffers) holds at submission; for device buffers a
    // null `ready` means the null stream. Returns the request id (>= 0).
    int64_t submit(int n, int count, const void* H, const void* S, double* W, void* U, int* info,
                   bool on_device, hipStream_t ready);
    // Admission policy and slice size (pencils; 0 = whole requests). Applies
    // to slices admitted from now on.
    void set_policy(Policy policy, int slice);
    // 1 done, 0 queued or running.
    int poll(int64_t id);
    // Block until the request is done; times in ms from submission to
    // completion (latency) and from admission to completion (service).
    void wait(int64_t id, double* latency_ms, double* service_ms);
    ServingStats stats();
    size_t arena_in_use();
    // Device footprint of a request staged through the arena.
    static size_t request_bytes(int n, int count);

private:
    struct Request {
        int64_t id = 0;
        int n = 0, count = 0;
        const void* H = nullptr;
        const void* S = nullptr;
        double* W = nullptr;
        void* U = nullptr;
        int* info = nullptr;
        bool on_device = false;
        hipEvent_t ready = nullptr;   // recorded on the caller's stream at submission
        size_t footprint = 0;
        size_t offset = SIZE_MAX;
        bool admitted = false;        // first slice started (arena block held)
        int dispatched = 0;           // pencils handed to slots so far
        int inflight = 0;             // slices running
        bool done = false;
        std::string error;
        std::chrono::steady_clock::time_point t_submit, t_admit, t_done;
    };
    struct Slot {
        hipStream_t stream = nullptr;
        hipEvent_t done = nullptr;
        std::map<int, std::unique_ptr<Solver>> solvers;   // order -> handle
        std::shared_ptr<Request> req;                      // request of the running slice
    };

    void loop_();
    // With mu_ held: start one slice on a free slot; false when no slot is
    // free, nothing is queued, or the chosen request's footprint does not fit.
    bool admit_one_();
    void launch_(Slot& s, Request& r, int off, int cnt);
    void finish_(const std::shared_ptr<Request>& r);      // with mu_ held

    int device_ = 0;
    Mode mode_ = Mode::Unified;
    int max_count_ = 0;
    std::vector<int> orders_;
    std::vector<Slot> slots_;
    void* dev_base_ = nullptr;
    std::unique_ptr<Arena> arena_;

    std::mutex mu_;
    std::condition_variable cv_done_;
    std::condition_variable cv_work_;
    std::deque<std::shared_ptr<Request>> pending_;   // requests with slices left, arrival order
    std::vector<hipEvent_t> event_pool_;
    Policy policy_ = Policy::Fifo;
    int slice_ = 0;
    std::unordered_map<int64_t, std::shared_ptr<Request>> requests_;
    int64_t next_id_ = 0;
    ServingStats stats_;
    bool stop_ = false;
    std::thread worker_;
};

}  // namespace sgeig
