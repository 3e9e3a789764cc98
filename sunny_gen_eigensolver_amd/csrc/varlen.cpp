#include "varlen.hpp"

#include "solver.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>

namespace sgeig {

namespace {

constexpr int kMaxVarlenOrder = Solver::kMaxOrder;
size_t al256(size_t x) { return (x + 255) / 256 * 256; }

}  // namespace

VarlenSolver::VarlenSolver(int max_n, int device_id, const std::vector<int>& buckets,
                           CacheMode cache, Tuning tuning)
    : max_n_(max_n), device_(device_id) {
    if (max_n < 1 || max_n > kMaxVarlenOrder)
        throw std::invalid_argument("max_n must be in [1, " + std::to_string(kMaxVarlenOrder) + "]");
    if (buckets.empty()) {
        for (int o = 4; o < max_n; o += (o < 64 ? 4 : o < 128 ? 8 : 16)) orders_.push_back(o);
        orders_.push_back(max_n);
    } else {
        orders_ = buckets;
        for (size_t i = 0; i < orders_.size(); ++i) {
            if (orders_[i] < 1 || orders_[i] > kMaxVarlenOrder)
                throw std::invalid_argument("bucket order out of range [1, " +
                                            std::to_string(kMaxVarlenOrder) + "]");
            if (i && orders_[i] <= orders_[i - 1])
                throw std::invalid_argument("bucket orders must be strictly ascending");
        }
        if (orders_.back() < max_n)
            throw std::invalid_argument("largest bucket order is below max_n");
    }
    int count = 0;
    SG_HIP(hipGetDeviceCount(&count));
    if (device_id < 0 || device_id >= count)
        throw std::invalid_argument("device_id " + std::to_string(device_id) + " out of range");
    SG_HIP(hipSetDevice(device_id));
    for (int o : orders_) {
        Bucket b;
        b.order = o;
        // The FP64 split mode at the bucket's order: with Tuning::Autotune its
        // own timed sweep (workgroup size; LDS-resident or staged from order 48).
        b.solver = std::make_unique<Solver>(o, device_id, 0, tuning, cache, Mode::Hybrid);
        b.nw = b.solver->staged() ? 1 : b.solver->block_dim_x() / 64;
        b.gather = b.solver->function("geneig_lg_gather");
        b.scatter = b.solver->function("geneig_lg_scatter");
        const size_t per = 3 * (size_t)o * o * 16;
        b.round = b.solver->staged()
                      ? kLargeRound
                      : (int)std::min<size_t>(kMaxRound, std::max<size_t>(kLargeRound, kRoundBytes / per));
        buckets_.push_back(std::move(b));
    }
    SG_HIP(hipEventCreateWithFlags(&staged_, hipEventDisableTiming));
    SG_HIP(hipEventCreateWithFlags(&done_, hipEventDisableTiming));
    SG_HIP(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
    for (int k = 0; k < kBranches; ++k) {
        SG_HIP(hipStreamCreateWithFlags(&branch_[k], hipStreamNonBlocking));
        SG_HIP(hipEventCreateWithFlags(&join_[k], hipEventDisableTiming));
    }
    if (const char* v = std::getenv("SUNNY_GENEIG_GRAPH")) graphs_ = std::atoi(v) != 0;
}

VarlenSolver::~VarlenSolver() {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device_);
    if (done_) (void)hipEventSynchronize(done_);
    for (auto& e : graph_cache_) {
        (void)hipGraphExecDestroy(e.exec);
        (void)hipFree(e.desc);
    }
    graph_cache_.clear();
    if (cap_stream_) (void)hipStreamDestroy(cap_stream_);
    for (int k = 0; k < kBranches; ++k) {
        if (branch_[k]) (void)hipStreamSynchronize(branch_[k]);
        if (branch_[k]) (void)hipStreamDestroy(branch_[k]);
        if (join_[k]) (void)hipEventDestroy(join_[k]);
    }
    if (fork_) (void)hipEventDestroy(fork_);
    if (staged_) (void)hipEventDestroy(staged_);
    if (done_) (void)hipEventDestroy(done_);
    if (host_desc_) (void)hipHostFree(host_desc_);
    if (dev_desc_) (void)hipFree(dev_desc_);
    if (pad_) (void)hipFree(pad_);
    buckets_.clear();
    (void)hipSetDevice(cur);
}

void VarlenSolver::ensure_capacity_(int count) {
    if ((size_t)count <= desc_count_) return;
    // Earlier launches may still read the device descriptors.
    SG_HIP(hipEventSynchronize(done_));
    if (host_desc_) SG_HIP(hipHostFree(host_desc_));
    if (dev_desc_) SG_HIP(hipFree(dev_desc_));
    host_desc_ = dev_desc_ = nullptr;
    desc_count_ = 0;
    const size_t bytes = (size_t)count * 24;
    SG_HIP(hipHostMalloc(&host_desc_, bytes, hipHostMallocDefault));
    SG_HIP(hipMalloc(&dev_desc_, bytes));
    desc_count_ = (size_t)count;
}

namespace {
uint64_t orders_hash(const int* n, int count) {
    uint64_t h = 0xcbf29ce484222325ULL;
    const unsigned char* p = reinterpret_cast<const unsigned char*>(n);
    for (size_t i = 0; i < (size_t)count * sizeof(int); ++i) {
        h ^= p[i];
        h *= 0x100000001b3ULL;
    }
    return h;
}
}  // namespace

void VarlenSolver::clear_graphs_() {
    if (graph_cache_.empty()) return;
    SG_HIP(hipEventSynchronize(done_));
    for (auto& e : graph_cache_) {
        (void)hipGraphExecDestroy(e.exec);
        (void)hipFree(e.desc);
    }
    graph_cache_.clear();
}

void VarlenSolver::set_graphs(bool on) {
    graphs_ = on;
    if (!on) clear_graphs_();
}

void VarlenSolver::launch(const void* H, const void* S, double* W, void* U, int* info, const int* n,
                          int count, hipStream_t stream) {
    if (count < 0) throw std::invalid_argument("count < 0");
    if (count == 0) return;
    if (!n) throw std::invalid_argument("order array is NULL");
    const int nb = (int)buckets_.size();
    std::vector<int> bucket_of(count), per_bucket(nb, 0);
    for (int k = 0; k < count; ++k) {
        if (n[k] < 1 || n[k] > max_n_)
            throw std::invalid_argument("matrix " + std::to_string(k) + " has order " +
                                        std::to_string(n[k]) + " outside [1, " +
                                        std::to_string(max_n_) + "]");
        const int b = (int)(std::lower_bound(orders_.begin(), orders_.end(), n[k]) - orders_.begin());
        bucket_of[k] = b;
        ++per_bucket[b];
    }
    int cur = 0;
    SG_HIP(hipGetDevice(&cur));
    if (cur != device_) SG_HIP(hipSetDevice(device_));
    const uint64_t hash = graphs_ ? orders_hash(n, count) : 0;
    if (graphs_) {
        for (auto& e : graph_cache_) {
            if (e.H == H && e.S == S && e.W == W && e.U == U && e.info == info && e.count == count &&
                e.hash == hash) {
                ++graph_replays_;
                SG_HIP(hipStreamWaitEvent(stream, done_, 0));
                SG_HIP(hipGraphLaunch(e.exec, stream));
                SG_HIP(hipEventRecord(done_, stream));
                if (cur != device_) SG_HIP(hipSetDevice(cur));
                return;
            }
        }
    }

    ensure_capacity_(count);
    size_t pad_need = 0;
    bool grew = false;
    if (reserved_.size() != (size_t)nb) reserved_.assign(nb, 0);
    for (int b = 0; b < nb; ++b) {
        if (!per_bucket[b]) continue;
        const Bucket& B = buckets_[b];
        const size_t no = (size_t)B.order;
        // Room for at least kMinReserve matrices (at most a round), grown by
        // doubling: a later call with a few more pencils in a bucket then
        // seldom moves a workspace that captured graphs hold.
        const size_t round = (size_t)B.round;
        size_t r = std::min<size_t>(per_bucket[b], round);
        if (r > reserved_[b]) r = std::min(round, std::max({r, 2 * reserved_[b], (size_t)kMinReserve}));
        else r = reserved_[b];
        pad_need = std::max(pad_need, 3 * al256(r * no * no * 16) + al256(r * no * 8) + al256(r * 4));
        if (r > reserved_[b]) {
            if (reserved_[b]) grew = true;      // its workspace may move
            B.solver->reserve(r);
            reserved_[b] = r;
        }
    }
    if (pad_need > pad_bytes_) {
        SG_HIP(hipEventSynchronize(done_));
        if (pad_) SG_HIP(hipFree(pad_));
        pad_ = nullptr;
        pad_bytes_ = 0;
        SG_HIP(hipMalloc(&pad_, pad_need * kBranches));
        pad_bytes_ = pad_need;
        grew = true;
    }
    if (grew) clear_graphs_();               // captured grids hold the old areas

    // The previous upload must have left the pinned staging area.
    if (staged_pending_) SG_HIP(hipEventSynchronize(staged_));
    staged_pending_ = false;
    int64_t* moff = static_cast<int64_t*>(host_desc_);
    int64_t* woff = moff + count;
    int* ids = reinterpret_cast<int*>(woff + count);
    int* nn = ids + count;
    int64_t mo = 0, wo = 0;
    for (int k = 0; k < count; ++k) {
        moff[k] = mo;
        woff[k] = wo;
        nn[k] = n[k];
        mo += (int64_t)n[k] * n[k];
        wo += n[k];
    }
    // Counting sort of matrix ids by bucket (stable).
    std::vector<int> start(nb + 1, 0);
    for (int b = 0; b < nb; ++b) start[b + 1] = start[b] + per_bucket[b];
    {
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (int k = 0; k < count; ++k) ids[fill[bucket_of[k]]++] = k;
    }
    const size_t desc_bytes = (size_t)count * 24;
    if (!graphs_) {
        // Descriptors and scratch are reused: order this call after the last
        // one even when the caller switches streams.
        SG_HIP(hipStreamWaitEvent(stream, done_, 0));
        SG_HIP(hipMemcpyAsync(dev_desc_, host_desc_, desc_bytes, hipMemcpyHostToDevice, stream));
        SG_HIP(hipEventRecord(staged_, stream));
        staged_pending_ = true;
        enqueue_(H, S, W, U, info, dev_desc_, count, per_bucket, start, stream);
        SG_HIP(hipEventRecord(done_, stream));
        if (cur != device_) SG_HIP(hipSetDevice(cur));
        return;
    }
    // Capture: the graph gets its own descriptors (a later call with other
    // orders must not overwrite what it reads).
    if (graph_cache_.size() >= 4) clear_graphs_();
    void* desc = nullptr;
    SG_HIP(hipMalloc(&desc, desc_bytes));
    try {
        SG_HIP(hipMemcpy(desc, host_desc_, desc_bytes, hipMemcpyHostToDevice));
        if (!cap_stream_) SG_HIP(hipStreamCreateWithFlags(&cap_stream_, hipStreamNonBlocking));
        SG_HIP(hipStreamBeginCapture(cap_stream_, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_(H, S, W, U, info, desc, count, per_bucket, start, cap_stream_);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(cap_stream_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hipGraph_t g = nullptr;
        SG_HIP(hipStreamEndCapture(cap_stream_, &g));
        hipGraphExec_t ex = nullptr;
        const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        SG_HIP(ie);
        graph_cache_.push_back({H, S, W, U, info, count, hash, desc, ex});
        ++graph_captures_;
    } catch (...) {
        (void)hipFree(desc);
        if (cur != device_) (void)hipSetDevice(cur);
        throw;
    }
    SG_HIP(hipStreamWaitEvent(stream, done_, 0));
    SG_HIP(hipGraphLaunch(graph_cache_.back().exec, stream));
    SG_HIP(hipEventRecord(done_, stream));
    if (cur != device_) SG_HIP(hipSetDevice(cur));
}

void VarlenSolver::enqueue_(const void* H, const void* S, double* W, void* U, int* info, const void* desc,
                            int count, const std::vector<int>& per_bucket, const std::vector<int>& start,
                            hipStream_t stream) {
    const int nb = (int)buckets_.size();
    const int64_t* d_moff = static_cast<const int64_t*>(desc);
    const int64_t* d_woff = d_moff + count;
    const int* d_ids = reinterpret_cast<const int*>(d_woff + count);
    const int* d_nn = d_ids + count;
    // Buckets, largest orders first, each to the branch with the least work
    // so far (pencils x order^3); a branch runs its buckets' rounds one after
    // another through its own padded staging area. Every bucket has its own
    // solver, and so its own workspace.
    SG_HIP(hipEventRecord(fork_, stream));
    for (int k = 0; k < kBranches; ++k) SG_HIP(hipStreamWaitEvent(branch_[k], fork_, 0));
    double load[kBranches] = {};
    for (int b = nb - 1; b >= 0; --b) {
        if (!per_bucket[b]) continue;
        const Bucket& B = buckets_[b];
        const size_t no = (size_t)B.order;
        const int k = (int)(std::min_element(load, load + kBranches) - load);
        load[k] += (double)per_bucket[b] * (double)(no * no * no);
        hipStream_t bs = branch_[k];
        for (int off = 0; off < per_bucket[b]; off += B.round) {
            int cnt = std::min(per_bucket[b] - off, B.round);
            const size_t r = (size_t)cnt;
            char* base = static_cast<char*>(pad_) + (size_t)k * pad_bytes_;
            void* Hp = base;
            void* Sp = base + al256(r * no * no * 16);
            void* Up = base + 2 * al256(r * no * no * 16);
            double* Wp = reinterpret_cast<double*>(base + 3 * al256(r * no * no * 16));
            int* Ip = reinterpret_cast<int*>(base + 3 * al256(r * no * no * 16) + al256(r * no * 8));
            const int* list = d_ids + start[b] + off;
            void* ga[] = {&H, &S, &list, &d_moff, &d_nn, &cnt, &Hp, &Sp};
            SG_HIP(hipModuleLaunchKernel(B.gather, (unsigned)cnt, 1, 1, 256, 1, 1, 0, bs, ga, nullptr));
            B.solver->launch_unordered(Hp, Sp, Wp, Up, Ip, cnt, bs);
            void* sa[] = {&Wp, &Up, &Ip, &list, &d_moff, &d_woff, &d_nn, &cnt, &W, &U, &info};
            SG_HIP(hipModuleLaunchKernel(B.scatter, (unsigned)cnt, 1, 1, 256, 1, 1, 0, bs, sa, nullptr));
        }
    }
    for (int k = 0; k < kBranches; ++k) {
        SG_HIP(hipEventRecord(join_[k], branch_[k]));
        SG_HIP(hipStreamWaitEvent(stream, join_[k], 0));
    }
}

}  // namespace sgeig
