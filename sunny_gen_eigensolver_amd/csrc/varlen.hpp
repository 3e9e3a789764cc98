// Uneven batches: matrices of different orders solved by one call. Every
// matrix goes to the smallest bucket order >= its own and is embedded in the
// bucket's order as (H (+) pad I, S (+) I). The host sorts the work list by
// bucket, so each bucket reads a contiguous run of it. Each bucket gathers its
// pencils into a padded contiguous batch, runs its FP64 split-mode solver on
// it (the matrix-core pipeline up to order 64, the staged kernels above) and
// scatters the results back, in rounds.
//
// No counterpart in the reference (its solvers take one order per handle,
// src/unified/unified_solver.hpp:81-217).
#pragma once

#include "codeobj.hpp"
#include "solver.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

namespace sgeig {

class VarlenSolver {
public:
    // Buckets: ascending orders in [1, 192]; empty = multiples of 4 up to 64,
    // then of 8 up to 128, then of 16, ending with max_n itself. With
    // Tuning::Autotune every bucket takes what the split-mode solver's timed
    // sweep picks for its order (workgroup size, and from order 48 the staged
    // pipeline; persisted per order and architecture); Suggested takes the
    // fixed rule. Buckets above 64 run the staged kernels (one size).
    VarlenSolver(int max_n, int device_id, const std::vector<int>& buckets, CacheMode cache,
                 Tuning tuning = Tuning::Suggested);
    ~VarlenSolver();
    VarlenSolver(const VarlenSolver&) = delete;
    VarlenSolver& operator=(const VarlenSolver&) = delete;

    // Matrices packed back to back: matrix k (order n[k]) at element offset
    // sum_{j<k} n[j]^2 of H, S and U, its eigenvalues at sum_{j<k} n[j] of W,
    // its status at info[k]. Asynchronous on `stream`; `n` is read before
    // return. Calls are ordered after the handle's previous call whatever the
    // stream. The first call with given buffers and orders captures every
    // grid of the call (each bucket's gather, solve and scatter rounds) into
    // one hipGraph with its own copy of the
    // work-list descriptors; a later call with the same buffers and orders
    // (compared by count and a 64-bit hash of the orders) replays it: one
    // graph launch, no host sort, no upload. SUNNY_GENEIG_GRAPH=0 at creation
    // or set_graphs(false) launches the grids directly.
    void launch(const void* H, const void* S, double* W, void* U, int* info, const int* n,
                int count, hipStream_t stream);
    void set_graphs(bool on);
    long long graph_captures() const { return graph_captures_; }
    long long graph_replays() const { return graph_replays_; }

    int max_n() const { return max_n_; }
    const std::vector<int>& buckets() const { return orders_; }
    // Matrices per workgroup of each bucket's QL grids (one matrix per wave;
    // 1 for a bucket on the staged kernels).
    int bucket_waves(int i) const { return buckets_[i].nw; }

    // Matrices per round of a bucket on the staged kernels.
    static constexpr int kLargeRound = 4096;
    // Rounds of the other buckets: as many matrices as keep the padded H, S
    // and U of a round within kRoundBytes, at least kLargeRound and at most
    // kMaxRound.
    static constexpr size_t kRoundBytes = size_t(1) << 30;
    // Buckets run on this many branches of the call (streams forked from and
    // joined back into the caller's), each with its own padded staging area,
    // so that one bucket's kernels fill the CUs another's leave idle.
    static constexpr int kBranches = 4;
    static constexpr int kMaxRound = 65536;
    // Matrices a bucket's workspace holds at least once the bucket is used.
    static constexpr int kMinReserve = 256;

private:
    struct Bucket {
        int order = 0;
        int nw = 0;
        int round = 0;                      // matrices per round
        // The bucket's FP64 split-mode solver and its gather / scatter.
        std::unique_ptr<Solver> solver;
        hipFunction_t gather = nullptr, scatter = nullptr;
    };
    void ensure_capacity_(int count);
    // Every grid of one call, on `stream`, reading the work-list descriptors
    // at `desc` (moff | woff | ids | nn for `count` matrices).
    void enqueue_(const void* H, const void* S, double* W, void* U, int* info, const void* desc, int count,
                  const std::vector<int>& per_bucket, const std::vector<int>& start, hipStream_t stream);

    struct GraphEntry {
        const void* H;
        const void* S;
        double* W;
        void* U;
        int* info;
        int count;
        uint64_t hash;
        void* desc;              // the descriptors this graph reads (device)
        hipGraphExec_t exec;
    };
    bool graphs_ = true;
    hipStream_t cap_stream_ = nullptr;
    std::vector<GraphEntry> graph_cache_;
    long long graph_captures_ = 0, graph_replays_ = 0;
    std::vector<size_t> reserved_;         // matrices each bucket's solver holds
    void clear_graphs_();

    int max_n_ = 0;
    int device_ = 0;
    std::vector<int> orders_;
    std::vector<Bucket> buckets_;
    // Descriptor staging: ids | nn (int32) then moff | woff (int64), pinned on
    // the host and mirrored on the device.
    void* host_desc_ = nullptr;
    void* dev_desc_ = nullptr;
    size_t desc_count_ = 0;
    hipEvent_t staged_ = nullptr;   // last upload from host_desc_ has finished
    bool staged_pending_ = false;
    hipEvent_t done_ = nullptr;     // last launch's grids have finished
    // Branch streams (default priority) and their fork / join events.
    hipStream_t branch_[kBranches] = {};
    hipEvent_t fork_ = nullptr;
    hipEvent_t join_[kBranches] = {};
    // Padded batches of the branches' rounds: H, S, U, W, info; pad_bytes_
    // per branch.
    void* pad_ = nullptr;
    size_t pad_bytes_ = 0;
};

}  // namespace sgeig
