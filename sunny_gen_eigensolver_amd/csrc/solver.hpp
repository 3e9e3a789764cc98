// Host runtime of the batched generalized eigensolver: the fused ("unified")
// and split ("hybrid") modes behind the C ABI.
//
// Reference counterparts: UnifiedGeneigSolver (src/unified/unified_solver.hpp:81-217)
// and HybridGeneigSolver (src/hybrid/hybrid_solver.hpp:34-89).
#pragma once

#include "codeobj.hpp"

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

namespace sgeig {

enum class Tuning { Suggested, Autotune };
enum class Mode { Unified, Hybrid, Mixed };

class Solver {
public:
    Solver(int n, int device_id, int bpb_request, Tuning tuning, CacheMode cache, Mode mode);
    ~Solver();
    Solver(const Solver&) = delete;
    Solver& operator=(const Solver&) = delete;

    // Asynchronous on `stream`; batch == 0 is a no-op. In the modes with a
    // workspace (split, mixed, staged orders) a launch is ordered after the
    // handle's previous launch even when the caller switches streams.
    void launch(const void* H, const void* S, double* W, void* U, int* info, int batch,
                hipStream_t stream);

    // As launch() for a real diagonal H: row b * hdiag_stride of Hd (FP64, n
    // entries; stride 0 = one diagonal for the whole batch, e.g. Sunny's
    // paraunitary metric diag(I, -I)). The mixed mode at orders up to 64 forms
    // M = W diag(h) W^H directly without H; the other paths expand the diagonal
    // into a full H in a workspace first.
    void launch_hdiag(const double* Hd, long long hdiag_stride, const void* S, double* W, void* U, int* info,
                      int batch, hipStream_t stream);

    // As launch(), but without ordering against this handle's other launches
    // and without graph capture, so that a caller which orders its launches
    // itself can issue it into a stream it is capturing (the uneven-batch
    // path). The workspace must already hold `batch` matrices (reserve()).
    void launch_unordered(const void* H, const void* S, double* W, void* U, int* info, int batch,
                          hipStream_t stream);

    // Contiguous chunks of `chunk` matrices launched round-robin over
    // `streams` (reference: UnifiedGeneigSolver::launch_chunked,
    // unified_solver.cpp:1078-1107). Fused-mode chunks run concurrently;
    // chunks of the modes with a workspace run in submission order.
    void launch_chunked(const void* H, const void* S, double* W, void* U, int* info, int total,
                        int chunk, const std::vector<hipStream_t>& streams);

    int n() const { return n_; }
    int device() const { return device_; }
    // Matrices per workgroup of the fused kernel; the split and mixed modes
    // report 1, as the reference's split mode does (c_api.cpp:41-49).
    int batches_per_block() const { return (mode_ == Mode::Unified) ? bpb_ : 1; }
    int block_dim_x() const { return block_dim_; }
    unsigned shared_mem_bytes() const { return lds_bytes_; }
    Mode mode() const { return mode_; }
    const char* cache_layer() const { return layer_name(layer_); }
    double acquire_ms() const { return acquire_ms_; }
    bool was_autotuned() const { return autotuned_; }
    // HBM-staged pipeline (always above order 64; below it when the autotuner
    // or SUNNY_GENEIG_STAGED chose it) rather than the LDS-resident kernels.
    bool staged() const { return large_; }
    // The autotuner weighs the staged pipeline from this order up.
    static constexpr int kMinStagedTuneOrder = 48;

    // Suggested workgroup size for order n (a model, no measurement).
    static int suggested_block_dim(int n);

    // Same results as launch(), with the mode's device workspace carved from
    // the caller's `scratch` (`scratch_bytes` long, 256-byte aligned) instead
    // of the handle's own: no allocation, no graph capture, QL without the
    // rotation log. Throws std::invalid_argument when the batch needs more
    // than scratch_bytes (the staged orders run in rounds that fit instead).
    void launch_in_scratch(const void* H, const void* S, double* W, void* U, int* info, int batch,
                           hipStream_t stream, void* scratch, size_t scratch_bytes);
    // Scratch bytes launch_in_scratch needs for `batch` matrices (for the
    // staged orders: one matrix's area, the least a round needs).
    static size_t scratch_bytes_needed(Mode mode, int n, size_t batch);

private:
    // Device workspace of one launch: the split mode's off-diagonals and Z,
    // the mixed mode's hand-offs, the QL rotation logs (null: QL without the
    // log) and the staged kernels' per-matrix areas with the matrices one
    // round may hold.
    struct Workspace {
        void* split = nullptr;
        void* mixed = nullptr;
        void* qlog = nullptr;
        void* staged = nullptr;
        size_t staged_batch = 0;
        // A diagonal H (launch_hdiag): its diagonals, their stride, and the
        // full H it is expanded into for the paths that read H (null: formed
        // from the diagonal directly).
        const double* hdiag = nullptr;
        long long hdiag_stride = 0;
        void* hfull = nullptr;
    };
    void launch_impl_(const void* H, const double* Hd, long long hds, const void* S, double* W, void* U,
                      int* info, int batch, hipStream_t stream);
    bool hdiag_direct_() const { return (mode_ == Mode::Mixed || mode_ == Mode::Hybrid) && !large_; }
    void ensure_hbuf_(size_t batch);
    void* hbuf_ = nullptr;             // full H expanded from a diagonal
    size_t hbuf_batch_ = 0;
    hipFunction_t k_expand_ = nullptr, k_mx_formm_diag_ = nullptr;
    Workspace own_workspace_(size_t batch);   // grows the handle's workspace as needed

    void bind_module_(int block_dim, CacheMode cache);
    // (block_dim, staged) chosen by the timed sweep, or read from its table.
    std::pair<int, int> autotune_(CacheMode cache);
    void launch_unified_(const void* H, const void* S, double* W, void* U, int* info, int batch,
                         hipStream_t stream);
    void launch_hybrid_(const void* H, const void* S, double* W, void* U, int* info, int batch,
                        hipStream_t stream, const Workspace& ws);
    void launch_mixed_(const void* H, const void* S, double* W, void* U, int* info, int batch,
                       hipStream_t stream, const Workspace& ws);
    void launch_large_(const void* H, const void* S, double* W, void* U, int* info, int batch,
                       hipStream_t stream, const Workspace& ws);
    void ensure_workspace_(size_t batch);
    void ensure_mixed_workspace_(size_t batch);
    void ensure_large_workspace_(size_t batch);
    void ensure_qlog_(size_t batch);

    int n_ = 0;
    int device_ = 0;
    int bpb_ = 1;            // matrices per workgroup = waves per workgroup
    int block_dim_ = 0;
    unsigned lds_bytes_ = 0;
    Mode mode_ = Mode::Unified;
    CacheLayer layer_ = CacheLayer::Compile;
    double acquire_ms_ = 0.0;
    bool autotuned_ = false;
    std::string arch_;
    std::shared_ptr<LoadedModule> module_;
    hipFunction_t fused_ = nullptr;
    hipFunction_t k_heev_ = nullptr;
    hipFunction_t k_qlfull_ = nullptr;     // QL of the matrices whose log overflowed
    // FP64 split mode: tridiagonalisation, V = Q Z and U = W^H V (with the
    // Cholesky and M = W H W^H it shares with the mixed mode, below).
    hipFunction_t k_front64_ = nullptr;
    hipFunction_t k_back64_ = nullptr;
    hipFunction_t k_backtri_ = nullptr;
    // QL recurrence recorder (64 matrices per wave) of the split QL and its
    // per-matrix rotation logs (split and mixed modes, orders up to 64).
    hipFunction_t k_qlrec_ = nullptr;
    void* qlog_ = nullptr;
    size_t qlog_batch_ = 0;
    // Split-mode workspace: W = L^-1, packed L and M, tridiagonal
    // off-diagonals and Z, one set per matrix.
    void* lbuf_ = nullptr;
    size_t lbuf_batch_ = 0;
    // Mixed mode: the FP32 module and the FP32 hand-offs between its stages
    // (L, stash then eigenvectors, eigenvalues, off-diagonals, Z), one set per
    // matrix; refinement steps and the relative closeness threshold.
    std::shared_ptr<LoadedModule> module32_;
    hipFunction_t k_mx_chol_ = nullptr;
    hipFunction_t k_mx_front_ = nullptr;
    hipFunction_t k_ql32_ = nullptr;
    hipFunction_t k_qlfull32_ = nullptr;
    hipFunction_t k_mx_formm_ = nullptr;
    bool w_packed_ = false;          // geneig_lg_chol leaves W packed (split and mixed modes)
    hipFunction_t k_back32_ = nullptr;
    hipFunction_t k_refine_ = nullptr;
    void* mbuf_ = nullptr;
    size_t mbuf_batch_ = 0;
    // Orders above 64 (the staged kernels, every mode): per-matrix workspace
    // for up to large_chunk(n) matrices at a time.
    bool large_ = false;     // staged kernels: order above 64, or staged_
    bool staged_ = false;    // staged kernels chosen at an order up to 64
    // k_lg_qll_: the lane-parallel QL recorder (orders up to 128; nullptr
    // above), used from lanes_min_batch_ matrices per launch: the measured
    // crossovers with the one-wave form at orders 96 and 128 (about 3,500
    // matrices in FP64, 9,000 in FP32; its time is nearly constant, the
    // one-wave form's grows with the batch). SUNNY_GENEIG_QL_LANES_MIN_BATCH
    // overrides it.
    hipFunction_t k_lg_qll_ = nullptr;
    int lg_qll_lanes_ = 64;       // its matrices per wave (the module states it)
    unsigned lg_hetrd_threads_ = 256;   // geneig_lg_hetrd's workgroup (the module states it)
    int ql_rec_lanes_ = 64;       // geneig_ql_record's matrices per wave
    static constexpr int kLanesMinBatch64 = 3584;
    static constexpr int kLanesMinBatch32 = 9216;
    int lanes_min_batch_ = kLanesMinBatch64;
    hipFunction_t k_lg_chol_ = nullptr, k_lg_formm_ = nullptr, k_lg_hetrd_ = nullptr, k_lg_ql_ = nullptr,
                  k_lg_qlz_ = nullptr, k_lg_qlzi_ = nullptr, k_lg_back_ = nullptr, k_lg_backtri_ = nullptr, k_lg_refine_ = nullptr;
    unsigned lg_back_threads_ = 256;   // workgroup size of geneig_lg_back (its launch bound)
    void* gbuf_ = nullptr;
    size_t gbuf_batch_ = 0;
    int refine_iters_ = 2;
    double refine_delta_ = 1e-4;
    // Largest entry the last refinement step's correction may have; above it
    // info = N. Measured on the benchmark generator (N = 54, 96, 192) and the
    // fixtures: converged matrices stay below 1e-6 (a tail of pairs with
    // gaps just above the closeness threshold reaches 1e-7), so 1e-5 leaves
    // a factor 10. SUNNY_GENEIG_REFINE_TOL overrides it.
    double refine_tol_ = 1e-5;
    // Sweeps a QL log may hold before the matrix is redone without a log (the
    // kernels clamp it to the log's capacity); SUNNY_GENEIG_QL_SWEEP_CAP lowers
    // it, which tests use to run the no-log path.
    int ql_sweep_cap_ = 1 << 30;
    // Batches from this size on run QL through the log (recorder + applier);
    // smaller ones run the no-log kernel for every matrix, since the recorder
    // costs about one lane's whole recurrence whatever the batch (N = 54:
    // 1.9 ms in FP64, 1.3 ms in FP32). Measured crossovers at orders 24..64:
    // about 6,000 matrices in FP64 (split mode), 10,000-12,000 in FP32
    // (mixed mode). SUNNY_GENEIG_QL_LOG_MIN_BATCH overrides it.
    static constexpr int kQlLogMinBatch64 = 6144;
    static constexpr int kQlLogMinBatch32 = 12288;
    static int ql_log_min_batch_default(Mode mode) {
        return mode == Mode::Mixed ? kQlLogMinBatch32 : kQlLogMinBatch64;
    }
    int ql_log_min_batch_ = kQlLogMinBatch64;
    bool last_ql_log_ = true;               // QL path of the last split / mixed launch

    // Multi-kernel launches (split, mixed, staged) are captured once per
    // (buffers, batch) into a hipGraph on an internal stream and replayed on
    // the caller's stream afterwards; SUNNY_GENEIG_GRAPH=0 or set_graphs(false)
    // launches the kernels directly.
    struct GraphEntry {
        const void* H;
        const double* hd;
        long long hds;
        const void* S;
        double* W;
        void* U;
        int* info;
        int batch;
        hipGraphExec_t exec;
    };
    bool graphs_ = true;
    hipStream_t cap_stream_ = nullptr;
    std::vector<GraphEntry> graph_cache_;
    long long graph_captures_ = 0, graph_replays_ = 0;
    void clear_graphs_();
    void dispatch_(const void* H, const void* S, double* W, void* U, int* info, int batch, hipStream_t stream,
                   const Workspace& ws);

    // Recorded after every launch of a mode with a workspace; the next launch
    // waits for it on its own stream.
    hipEvent_t done_ = nullptr;

    // Stage timing: when enabled, events bracket every kernel of a launch on
    // its stream; stage_times() reads the last launch's durations.
    bool timing_ = false;
    std::vector<hipEvent_t> ev_;
    std::vector<hipFunction_t> ev_fn_;      // kernel each event closes
    int ev_used_ = 0;
    hipFunction_t last_fn_ = nullptr;       // kernel launched last
    void mark_(hipStream_t stream);
    void launch_(hipFunction_t f, unsigned grid, unsigned block, hipStream_t stream, void** args);

public:
    int refine_iters() const { return refine_iters_; }
    // A kernel of this handle's FP64 module (the staged orders' uneven-batch
    // gather / scatter live there).
    hipFunction_t function(const char* name) const { return module_->get(name); }
    double refine_delta() const { return refine_delta_; }
    // Grow the mode's device workspace for `batch` matrices now (no-op for the
    // fused mode), so later launches of up to `batch` allocate nothing; returns
    // its size in bytes.
    size_t reserve(size_t batch);
    static size_t workspace_bytes(Mode mode, int n, size_t batch);
    // Orders above kMaxLdsOrder run the staged kernels; up to kMaxOrder.
    static constexpr int kMaxLdsOrder = 64;
    static constexpr int kMaxOrder = 192;
    // Staged kernels: workspace bytes per matrix, and matrices per launch
    // round (a larger batch is solved in rounds of this many).
    static size_t large_matrix_bytes(int n);
    static size_t large_chunk(int n);
    // Bytes of one matrix's QL rotation log (orders up to 64).
    static size_t ql_log_bytes(int n);
    void set_stage_timing(bool on);
    void set_graphs(bool on);
    long long graph_captures() const { return graph_captures_; }
    long long graph_replays() const { return graph_replays_; }
    // Fills ms[0 .. min(n, stages)) and returns the number of stages of the
    // last timed launch (0 if none); waits for that launch to finish.
    int stage_times(double* ms, int n);
    // Kernel name of stage i of the last timed launch ("" if none).
    const char* stage_label(int i) const;
    bool last_ql_log() const { return last_ql_log_; }
};

}  // namespace sgeig
