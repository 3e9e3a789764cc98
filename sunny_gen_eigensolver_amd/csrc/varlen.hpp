// Uneven batches: matrices of different orders solved by one call. Every
// matrix goes to the smallest bucket order >= its own and is embedded in the
// bucket's order as (H (+) pad I, S (+) I). The host sorts the work list by
// bucket, so each bucket reads a contiguous run of it. Buckets up to 64 are
// one grid of the LDS-resident geneig_varlen kernel each; a bucket above 64
// gathers its pencils into a padded contiguous batch, runs the staged FP64
// kernels on it and scatters the results back (in rounds of kLargeRound).
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
    // Tuning::Autotune every bucket up to 64 takes the workgroup size (and so
    // the matrices per workgroup) that the fused solver's timed sweep picks for
    // its order (persisted per order and architecture); Suggested takes the
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
    // grid of the call (the LDS-resident buckets, and the staged buckets'
    // gather, solve and scatter) into one hipGraph with its own copy of the
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
    // Matrices per workgroup of each bucket's grid.
    int bucket_waves(int i) const { return buckets_[i].nw; }

    // Workgroups per grid launch; longer runs are split into several grids
    // that reuse one scratch area.
    static constexpr int kMaxGrid = 4096;
    // Matrices per round of a bucket above 64.
    static constexpr int kLargeRound = 4096;

private:
    struct Bucket {
        int order = 0;
        int block_dim = 0;
        int nw = 0;
        std::shared_ptr<LoadedModule> module;
        hipFunction_t fn = nullptr;
        // Buckets above 64: the staged FP64 solver and its gather / scatter.
        std::unique_ptr<Solver> large;
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
    std::vector<size_t> large_reserved_;   // matrices each staged bucket's solver holds
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
    // Per-workgroup factor stash, sized for the largest bucket's grid.
    void* scratch_ = nullptr;
    size_t scratch_bytes_ = 0;
    // Padded batch of a large bucket's round: H, S, U, W, info.
    void* pad_ = nullptr;
    size_t pad_bytes_ = 0;
};

}  // namespace sgeig
