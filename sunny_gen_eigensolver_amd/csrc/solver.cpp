#include "solver.hpp"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <complex>
#include <cstdio>
#include <fstream>
#include <map>
#include <mutex>
#include <random>
#include <sstream>
#include <vector>

namespace sgeig {

namespace {

// Workspace of one round of the staged kernels (orders above 64).
constexpr size_t kLargeRoundBytes = size_t(32) << 30;

// Persisted autotune winners: "<n> <arch> <mode> <block_dim> <staged>" lines
// in <cache_dir>/autotune.tsv (mode 0 fused, 1 split, 2 mixed; staged 1 = the
// HBM-staged pipeline). The reference memoises per process only
// (unified_solver.cpp:647, :970-993); this table also survives restarts.
struct TuneKey {
    int n;
    std::string arch;
    int mode;
    bool operator<(const TuneKey& o) const {
        if (n != o.n) return n < o.n;
        if (mode != o.mode) return mode < o.mode;
        return arch < o.arch;
    }
};
std::mutex g_tune_mutex;
std::map<TuneKey, std::pair<int, int>> g_tune;
bool g_tune_loaded = false;

std::string tune_path() { return cache_dir() + "/autotune.tsv"; }

void load_tune_table() {
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    std::ifstream f(tune_path());
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream ls(line);
        TuneKey k;
        int bd = 0, staged = 0;
        if (ls >> k.n >> k.arch >> k.mode >> bd >> staged) g_tune[k] = {bd, staged};
    }
}

void save_tune_table() {
    const std::string path = tune_path();
    const std::string tmp = path + ".tmp";
    std::ofstream f(tmp, std::ios::trunc);
    if (!f) return;
    for (auto& kv : g_tune)
        f << kv.first.n << ' ' << kv.first.arch << ' ' << kv.first.mode << ' ' << kv.second.first << ' '
          << kv.second.second << '\n';
    f.close();
    std::rename(tmp.c_str(), path.c_str());
}

// Tuning workload: H = A + A^H and S = A A^H + n I from a seeded Gaussian,
// the same shape of problem as the reference's tuning generator
// (unified_solver.cpp:606-638).
void tune_inputs(int n, int batch, uint64_t seed, std::vector<std::complex<double>>& H,
                 std::vector<std::complex<double>>& S) {
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> g(0.0, 1.0);
    const size_t nn = (size_t)n * n;
    H.assign(nn * batch, 0.0);
    S.assign(nn * batch, 0.0);
    std::vector<std::complex<double>> A(nn);
    for (int b = 0; b < batch; ++b) {
        for (auto& a : A) a = {g(rng), g(rng)};
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                H[nn * b + (size_t)j * n + i] = A[(size_t)j * n + i] + std::conj(A[(size_t)i * n + j]);
        for (auto& a : A) a = {g(rng), g(rng)};
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                std::complex<double> acc = 0.0;
                for (int k = 0; k < n; ++k) acc += A[(size_t)k * n + i] * std::conj(A[(size_t)k * n + j]);
                if (i == j) acc += (double)n;
                S[nn * b + (size_t)j * n + i] = acc;
            }
    }
}

struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) { SG_HIP(hipMalloc(&p, bytes)); }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

int Solver::suggested_block_dim(int n) {
    // One matrix per workgroup; its LDS footprint allows one workgroup per CU
    // at the larger orders, so those get more waves for the dense stages.
    if (n <= 16) return 64;
    if (n <= 32) return 128;
    return 256;
}

namespace {
// Mixed-mode workspace for a batch of B matrices of order n (byte offsets):
// W = L^-1 (complex128; room for the full matrix, packed where geneig_lg_chol
// packs it), L packed (complex128), the complex64 stash / V, the FP32
// eigenvalues and off-diagonals, Z (FP32) and M's lower triangle packed
// (complex128).
struct MixedLayout {
    size_t wd, lp, v32, w32, e32, z32, mp, total;
};
// The FP64 split mode's workspace: W = L^-1 (room for the full matrix;
// packed where geneig_lg_chol packs it), L and M (packed complex128), the
// tridiagonal's off-diagonal and Z (float64).
struct HybridLayout {
    size_t wd, lp, mp, e, z, total;
};
HybridLayout hybrid_layout(size_t n, size_t B) {
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    HybridLayout h;
    h.wd = 0;
    h.lp = al(h.wd + B * n * n * 16);
    h.mp = al(h.lp + B * n * (n + 1) / 2 * 16);
    h.e = al(h.mp + B * n * (n + 1) / 2 * 16);
    h.z = al(h.e + B * n * 8);
    h.total = al(h.z + B * n * n * 8);
    return h;
}

MixedLayout mixed_layout(size_t n, size_t B) {
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    MixedLayout m;
    m.wd = 0;
    m.lp = al(m.wd + B * n * n * 16);
    m.v32 = al(m.lp + B * n * (n + 1) / 2 * 16);
    m.w32 = al(m.v32 + B * n * n * 8);
    m.e32 = al(m.w32 + B * n * 4);
    m.z32 = al(m.e32 + B * n * 4);
    m.mp = al(m.z32 + B * n * n * 4);
    m.total = al(m.mp + B * n * (n + 1) / 2 * 16);
    return m;
}
}  // namespace

namespace {
// An int the module states (a kernel's launch shape).
int module_int(LoadedModule& m, const char* name) {
    hipDeviceptr_t sym = nullptr;
    size_t bytes = 0;
    int v = 0;
    SG_HIP(hipModuleGetGlobal(&sym, &bytes, m.module, name));
    SG_HIP(hipMemcpyDtoH(&v, sym, sizeof(v)));
    if (v <= 0 || v > 1024) throw std::runtime_error(std::string("module states ") + name + " = " + std::to_string(v));
    return v;
}
}  // namespace

Solver::Solver(int n, int device_id, int bpb_request, Tuning tuning, CacheMode cache, Mode mode)
    : n_(n), device_(device_id), mode_(mode) {
    if (n < 1) throw std::invalid_argument("matrix size must be >= 1");
    if (n > kMaxOrder)
        throw std::invalid_argument("matrix size " + std::to_string(n) + " exceeds the supported maximum of " +
                                    std::to_string(kMaxOrder));
    large_ = n > kMaxLdsOrder;
    if (const char* v = std::getenv("SUNNY_GENEIG_GRAPH")) graphs_ = std::atoi(v) != 0;
    if (const char* v = std::getenv("SUNNY_GENEIG_QL_SWEEP_CAP")) ql_sweep_cap_ = std::max(0, std::atoi(v));
    ql_log_min_batch_ = ql_log_min_batch_default(mode);
    if (const char* v = std::getenv("SUNNY_GENEIG_QL_LOG_MIN_BATCH")) ql_log_min_batch_ = std::max(0, std::atoi(v));
    lanes_min_batch_ = mode == Mode::Mixed ? kLanesMinBatch32 : kLanesMinBatch64;
    if (const char* v = std::getenv("SUNNY_GENEIG_QL_LANES_MIN_BATCH")) lanes_min_batch_ = std::max(0, std::atoi(v));
    // One matrix per wave: a workgroup of k waves solves k matrices, and its
    // LDS work area and register budget allow k <= 4. The reference returns
    // NULL for an explicit request it cannot build (unified_solver.cpp:1021-1028);
    // so does this library, with this message.
    if (bpb_request > 4)
        throw std::invalid_argument("batches_per_block_request=" + std::to_string(bpb_request) +
                                    " is not supported: at most 4 matrices per workgroup (one wave "
                                    "per matrix, LDS work area shared by the workgroup); pass 1..4, "
                                    "or <= 0 to select automatically");
    int count = 0;
    SG_HIP(hipGetDeviceCount(&count));
    if (device_id < 0 || device_id >= count)
        throw std::invalid_argument("device_id " + std::to_string(device_id) + " out of range");
    SG_HIP(hipSetDevice(device_id));
    arch_ = device_arch(device_id);
    SG_HIP(hipEventCreateWithFlags(&done_, hipEventDisableTiming));
    // A workgroup of W waves solves W matrices (one QL chain per wave), so a
    // requested matrices-per-block count fixes the workgroup size.
    int bd = (bpb_request > 0) ? 64 * bpb_request : suggested_block_dim(n);
    // SUNNY_GENEIG_STAGED=1 runs the HBM-staged pipeline at orders up to 64
    // as well (the autotuner's other blocking; tests use it to check it).
    if (const char* v = std::getenv("SUNNY_GENEIG_STAGED"))
        if (std::atoi(v) != 0 && bpb_request <= 0) staged_ = true;
    if (tuning == Tuning::Autotune && bpb_request <= 0 && !large_ && !staged_) {
        const auto choice = autotune_(cache);
        bd = choice.first;
        staged_ = choice.second != 0;
        autotuned_ = true;
    }
    large_ = n > kMaxLdsOrder || staged_;
    // Staged kernels: one matrix per 256-thread workgroup.
    if (large_) bd = 256;
    bind_module_(bd, cache);
}

Solver::~Solver() {
    for (auto& e : ev_) (void)hipEventDestroy(e);
    if (done_) {
        (void)hipEventSynchronize(done_);
        (void)hipEventDestroy(done_);
    }
    clear_graphs_();
    if (cap_stream_) (void)hipStreamDestroy(cap_stream_);
    if (lbuf_ || mbuf_ || gbuf_ || qlog_ || hbuf_) {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(device_);
        if (lbuf_) (void)hipFree(lbuf_);
        if (mbuf_) (void)hipFree(mbuf_);
        if (gbuf_) (void)hipFree(gbuf_);
        if (qlog_) (void)hipFree(qlog_);
        if (hbuf_) (void)hipFree(hbuf_);
        (void)hipSetDevice(cur);
    }
}

void Solver::bind_module_(int block_dim, CacheMode cache) {
    clear_graphs_();                         // captured launches hold the old kernels
    ModuleKey key{n_, block_dim, arch_};
    key.staged = staged_ && n_ <= kMaxLdsOrder;
    module_ = acquire_module(key, cache, &layer_, &acquire_ms_);
    block_dim_ = block_dim;
    bpb_ = block_dim / 64;
    k_expand_ = module_->get("geneig_expand_hdiag");
    if (large_) {
        bpb_ = 1;
        k_lg_chol_ = module_->get("geneig_lg_chol");
        k_lg_formm_ = module_->get("geneig_lg_formm");
        k_lg_backtri_ = module_->get("geneig_lg_backtri");
        hipFunction_t lds_probe = nullptr;
        if (mode_ == Mode::Mixed) {
            ModuleKey k32 = key;
            k32.fp32 = true;
            CacheLayer l32;
            double ms32 = 0.0;
            module32_ = acquire_module(k32, cache, &l32, &ms32);
            acquire_ms_ += ms32;
            k_lg_hetrd_ = module32_->get("geneig_lg_hetrd");
            lg_hetrd_threads_ = (unsigned)module_int(*module32_, "geneig_lg_hetrd_threads");
            k_lg_ql_ = module32_->get("geneig_lg_qlr");
            k_lg_qll_ = module32_->find("geneig_lg_qlr_lanes");
            if (k_lg_qll_) lg_qll_lanes_ = module_int(*module32_, "geneig_lg_qlr_lanes_count");
            k_lg_qlz_ = module32_->get("geneig_lg_qlz");
            k_lg_qlzi_ = module32_->get("geneig_lg_qlz_inline");
            k_lg_back_ = module32_->get("geneig_lg_back");
            k_lg_refine_ = module_->get("geneig_lg_refine");
            if (const char* v = std::getenv("SUNNY_GENEIG_REFINE_ITERS")) refine_iters_ = std::max(0, std::atoi(v));
            if (const char* v = std::getenv("SUNNY_GENEIG_REFINE_DELTA")) refine_delta_ = std::atof(v);
            if (const char* v = std::getenv("SUNNY_GENEIG_REFINE_TOL")) refine_tol_ = std::atof(v);
        } else {
            k_lg_hetrd_ = module_->get("geneig_lg_hetrd");
            lg_hetrd_threads_ = (unsigned)module_int(*module_, "geneig_lg_hetrd_threads");
            k_lg_ql_ = module_->get("geneig_lg_qlr");
            k_lg_qll_ = module_->find("geneig_lg_qlr_lanes");
            if (k_lg_qll_) lg_qll_lanes_ = module_int(*module_, "geneig_lg_qlr_lanes_count");
            k_lg_qlz_ = module_->get("geneig_lg_qlz");
            k_lg_qlzi_ = module_->get("geneig_lg_qlz_inline");
            k_lg_back_ = module_->get("geneig_lg_back");
        }
        {
            hipDeviceptr_t sym = nullptr;
            size_t sym_bytes = 0;
            unsigned long long dev_ws = 0;
            SG_HIP(hipModuleGetGlobal(&sym, &sym_bytes, module_->module, "geneig_lg_ws_bytes"));
            SG_HIP(hipMemcpyDtoH(&dev_ws, sym, sizeof(dev_ws)));
            if (dev_ws != large_matrix_bytes(n_))
                throw std::runtime_error("staged workspace layout mismatch: module " + std::to_string(dev_ws) +
                                         " B, host " + std::to_string(large_matrix_bytes(n_)) + " B");
        }
        {
            // V = Q Z runs one wave per 16-column panel where it is built on
            // the matrix cores: the module states its workgroup size.
            hipDeviceptr_t sym = nullptr;
            size_t sym_bytes = 0;
            int nt = 0;
            LoadedModule* bm = (mode_ == Mode::Mixed) ? module32_.get() : module_.get();
            SG_HIP(hipModuleGetGlobal(&sym, &sym_bytes, bm->module, "geneig_lg_back_threads"));
            SG_HIP(hipMemcpyDtoH(&nt, sym, sizeof(nt)));
            if (nt <= 0 || nt % 64 != 0 || nt > 1024)
                throw std::runtime_error("geneig_lg_back states " + std::to_string(nt) + " threads per block");
            lg_back_threads_ = (unsigned)nt;
        }
        lds_probe = k_lg_hetrd_;
        int lds = 0;
        SG_HIP(hipFuncGetAttribute(&lds, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, lds_probe));
        lds_bytes_ = (unsigned)lds;
        return;
    }
    fused_ = module_->get("geneig_fused");
    k_heev_ = module_->get("geneig_split_ql");
    k_qlfull_ = module_->get("geneig_ql_full");
    k_qlrec_ = module_->get("geneig_ql_record");
    ql_rec_lanes_ = module_int(*module_, "geneig_ql_record_lanes");
    {
        hipDeviceptr_t sym = nullptr;
        size_t sym_bytes = 0;
        unsigned long long dev_q = 0;
        SG_HIP(hipModuleGetGlobal(&sym, &sym_bytes, module_->module, "geneig_ql_log_bytes"));
        SG_HIP(hipMemcpyDtoH(&dev_q, sym, sizeof(dev_q)));
        if (dev_q != ql_log_bytes(n_))
            throw std::runtime_error("QL log layout mismatch: module " + std::to_string(dev_q) + " B, host " +
                                     std::to_string(ql_log_bytes(n_)) + " B");
    }
    if (mode_ == Mode::Hybrid || mode_ == Mode::Mixed) {
        // The FP64 stages both modes share: Cholesky, L and W = L^-1, M = W H W^H.
        k_mx_chol_ = module_->get("geneig_lg_chol");
        k_mx_formm_ = module_->get("geneig_mx_formm");
        k_mx_formm_diag_ = module_->get("geneig_mx_formm_diag");
        // geneig_lg_chol leaves W packed (its lower triangle) where it fits LDS.
        hipDeviceptr_t sym = nullptr;
        size_t sym_bytes = 0;
        int packed = 0;
        SG_HIP(hipModuleGetGlobal(&sym, &sym_bytes, module_->module, "geneig_w_packed"));
        SG_HIP(hipMemcpyDtoH(&packed, sym, sizeof(packed)));
        w_packed_ = packed != 0;
    }
    if (mode_ == Mode::Hybrid) {
        k_front64_ = module_->get("geneig_mx_front");
        k_back64_ = module_->get("geneig_mx_back");
        k_backtri_ = module_->get("geneig_mx_backtri");
    }
    if (mode_ == Mode::Mixed) {
        k_refine_ = module_->get("geneig_mx_refine");
        ModuleKey k32 = key;
        k32.fp32 = true;
        CacheLayer l32;
        double ms32 = 0.0;
        module32_ = acquire_module(k32, cache, &l32, &ms32);
        acquire_ms_ += ms32;
        k_mx_front_ = module32_->get("geneig_mx_front");
        k_ql32_ = module32_->get("geneig_split_ql");
        k_qlfull32_ = module32_->get("geneig_ql_full");
        k_qlrec_ = module32_->get("geneig_ql_record");
        ql_rec_lanes_ = module_int(*module32_, "geneig_ql_record_lanes");
        k_back32_ = module32_->get("geneig_mx_back");
        if (const char* v = std::getenv("SUNNY_GENEIG_REFINE_ITERS")) refine_iters_ = std::max(0, std::atoi(v));
        if (const char* v = std::getenv("SUNNY_GENEIG_REFINE_DELTA")) refine_delta_ = std::atof(v);
        if (const char* v = std::getenv("SUNNY_GENEIG_REFINE_TOL")) refine_tol_ = std::atof(v);
    }
    // Reported LDS: the fused kernel's; in the split mode the largest of the
    // reduction kernels' (the reference reports its reduce kernel's,
    // c_api.cpp:41-49); in the mixed mode the largest of its kernels'.
    std::vector<hipFunction_t> fns = {fused_};
    if (mode_ == Mode::Hybrid) fns = {k_mx_chol_, k_mx_formm_, k_front64_};
    if (mode_ == Mode::Mixed) fns = {k_mx_chol_, k_mx_formm_, k_mx_front_, k_back32_, k_refine_};
    int max_lds = 0;
    SG_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device_));
    int lds = 0;
    for (hipFunction_t f : fns) {
        int l = 0;
        SG_HIP(hipFuncGetAttribute(&l, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, f));
        if (l > max_lds)
            throw std::runtime_error("kernel needs " + std::to_string(l) + " B of LDS, device allows " +
                                     std::to_string(max_lds));
        lds = std::max(lds, l);
    }
    lds_bytes_ = (unsigned)lds;
}

// Timed, correctness-gated sweep, generalising the reference's BlockDim sweep
// (unified_solver.cpp:649-925) to the choices this library has per order and
// mode: the workgroup size, which with one wave per matrix is also the number
// of matrices per workgroup (64 / 128 / 256 threads = 1 / 2 / 4 matrices),
// and, from order kMinStagedTuneOrder up, LDS-resident versus HBM-staged
// blocking (the staged pipeline: one matrix per 256-thread workgroup). Every
// candidate is compiled and run on the tuning batch; it must reproduce the
// suggested candidate's eigenvalues (1e-10 relative) and, on spot matrices,
// its eigenvectors (1 - |u^H S u_ref| < 1e-10 for eigenvalues separated by
// 1e-6 or more, the reference's checks at :858-885); the fastest valid one
// (median of 5 launches) wins and is persisted per (order, arch, mode).
std::pair<int, int> Solver::autotune_(CacheMode cache) {
    const TuneKey key{n_, arch_, (int)mode_};
    {
        std::lock_guard<std::mutex> lk(g_tune_mutex);
        load_tune_table();
        auto it = g_tune.find(key);
        if (it != g_tune.end()) return it->second;
    }
    const int suggested = suggested_block_dim(n_);
    std::vector<std::pair<int, int>> cands = {{suggested, 0}};
    for (int c : {64, 128, 256})
        if (c != suggested) cands.push_back({c, 0});
    if (n_ >= kMinStagedTuneOrder) cands.push_back({256, 1});
    const int batch = 1024;
    const size_t nn = (size_t)n_ * n_;
    std::vector<std::complex<double>> H, S;
    tune_inputs(n_, batch, 0xA5A5A5A5ULL ^ ((uint64_t)n_ << 16) ^ (uint64_t)mode_, H, S);
    DevBuf dH(nn * batch * 16), dS(nn * batch * 16), dU(nn * batch * 16), dW((size_t)n_ * batch * 8),
        dI((size_t)batch * 4);
    SG_HIP(hipMemcpy(dH.p, H.data(), nn * batch * 16, hipMemcpyHostToDevice));
    SG_HIP(hipMemcpy(dS.p, S.data(), nn * batch * 16, hipMemcpyHostToDevice));
    constexpr int kSpots = 8;
    std::vector<double> ref_w, w((size_t)n_ * batch);
    std::vector<std::complex<double>> ref_u, u(nn * kSpots);
    hipEvent_t e0, e1;
    SG_HIP(hipEventCreate(&e0));
    SG_HIP(hipEventCreate(&e1));
    std::pair<int, int> best = cands[0];
    double best_ms = 1e300;
    const bool graphs = graphs_;
    graphs_ = false;                         // direct launches: candidates change kernels
    auto run = [&]() {
        launch(dH.p, dS.p, (double*)dW.p, dU.p, (int*)dI.p, batch, nullptr);
    };
    // Spot matrices spread over the batch.
    auto spot = [&](int s) { return (size_t)s * (batch / kSpots); };
    for (const auto& c : cands) {
        try {
            staged_ = c.second != 0;
            large_ = n_ > kMaxLdsOrder || staged_;
            bind_module_(staged_ ? 256 : c.first, cache);
            run();
            SG_HIP(hipDeviceSynchronize());
        } catch (const std::exception& ex) {
            std::fprintf(stderr, "[sunny_geneig autotune] N=%d wg=%d %s skipped: %s\n", n_, c.first,
                         c.second ? "staged" : "lds", ex.what());
            continue;
        }
        SG_HIP(hipMemcpy(w.data(), dW.p, w.size() * 8, hipMemcpyDeviceToHost));
        for (int sp = 0; sp < kSpots; ++sp)
            SG_HIP(hipMemcpy(u.data() + nn * sp, static_cast<char*>(dU.p) + spot(sp) * nn * 16, nn * 16,
                             hipMemcpyDeviceToHost));
        if (ref_w.empty()) {
            ref_w = w;
            ref_u = u;
        } else {
            double worst = 0.0, worst_phase = 0.0;
            for (size_t i = 0; i < w.size(); ++i)
                worst = std::max(worst, std::fabs(w[i] - ref_w[i]) / std::max(std::fabs(ref_w[i]), 1.0));
            std::vector<std::complex<double>> Su(n_);
            for (int sp = 0; sp < kSpots; ++sp) {
                const std::complex<double>* Sb = S.data() + nn * spot(sp);
                const double* wr = ref_w.data() + (size_t)n_ * spot(sp);
                for (int col = 0; col < n_; ++col) {
                    const double gap = std::min(col > 0 ? wr[col] - wr[col - 1] : 1e300,
                                                col + 1 < n_ ? wr[col + 1] - wr[col] : 1e300);
                    if (gap < 1e-6) continue;
                    const std::complex<double>* ur = ref_u.data() + nn * sp + (size_t)col * n_;
                    const std::complex<double>* uc = u.data() + nn * sp + (size_t)col * n_;
                    for (int i = 0; i < n_; ++i) {
                        std::complex<double> acc = 0.0;
                        for (int k = 0; k < n_; ++k) acc += Sb[(size_t)k * n_ + i] * ur[k];
                        Su[i] = acc;
                    }
                    std::complex<double> inner = 0.0;
                    for (int i = 0; i < n_; ++i) inner += std::conj(uc[i]) * Su[i];
                    worst_phase = std::max(worst_phase, 1.0 - std::abs(inner));
                }
            }
            if (!(worst < 1e-10) || !(worst_phase < 1e-10)) {
                std::fprintf(stderr,
                             "[sunny_geneig autotune] N=%d wg=%d %s rejected (eig diff %.3e, phase %.3e)\n", n_,
                             c.first, c.second ? "staged" : "lds", worst, worst_phase);
                continue;
            }
        }
        std::vector<float> samples;
        for (int sm = 0; sm < 5; ++sm) {
            SG_HIP(hipEventRecord(e0, nullptr));
            run();
            SG_HIP(hipEventRecord(e1, nullptr));
            SG_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            SG_HIP(hipEventElapsedTime(&ms, e0, e1));
            samples.push_back(ms);
        }
        std::sort(samples.begin(), samples.end());
        const double med = samples[2];
        std::fprintf(stderr, "[sunny_geneig autotune] N=%d mode=%d wg=%d %s median %.3f ms (%.0f solves/s)\n", n_,
                     (int)mode_, c.first, c.second ? "staged" : "lds", med, batch / (med * 1e-3));
        if (med < best_ms) {
            best_ms = med;
            best = c;
        }
    }
    graphs_ = graphs;
    staged_ = false;
    large_ = n_ > kMaxLdsOrder;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    {
        std::lock_guard<std::mutex> lk(g_tune_mutex);
        g_tune[key] = best;
        if (cache == CacheMode::Auto) save_tune_table();
    }
    return best;
}

void Solver::launch(const void* H, const void* S, double* W, void* U, int* info, int batch,
                    hipStream_t stream) {
    launch_impl_(H, nullptr, 0, S, W, U, info, batch, stream);
}

void Solver::launch_hdiag(const double* Hd, long long hdiag_stride, const void* S, double* W, void* U, int* info,
                          int batch, hipStream_t stream) {
    if (hdiag_stride < 0) throw std::invalid_argument("hdiag_stride < 0");
    if (batch > 0 && !Hd) throw std::invalid_argument("diagonal of H is NULL");
    launch_impl_(nullptr, Hd, hdiag_stride, S, W, U, info, batch, stream);
}

void Solver::launch_impl_(const void* H, const double* Hd, long long hds, const void* S, double* W, void* U,
                          int* info, int batch, hipStream_t stream) {
    if (batch < 0) throw std::invalid_argument("batch_size < 0");
    if (batch == 0) return;
    int cur = 0;
    SG_HIP(hipGetDevice(&cur));
    if (cur != device_) SG_HIP(hipSetDevice(device_));
    ev_used_ = 0;
    const bool multi = large_ || mode_ != Mode::Unified || Hd;
    // The workspace is shared by every launch of this handle: order this one
    // after the previous one, whatever stream that ran on.
    if (multi) SG_HIP(hipStreamWaitEvent(stream, done_, 0));
    Workspace ws = multi ? own_workspace_((size_t)batch) : Workspace{};   // allocates outside capture
    if (Hd) {
        ws.hdiag = Hd;
        ws.hdiag_stride = hds;
        if (!hdiag_direct_()) {
            ensure_hbuf_((size_t)batch);
            ws.hfull = hbuf_;
        }
    }
    if (graphs_ && multi && !timing_) {
        GraphEntry* hit = nullptr;
        for (auto& e : graph_cache_)
            if (e.H == H && e.hd == Hd && e.hds == hds && e.S == S && e.W == W && e.U == U && e.info == info &&
                e.batch == batch)
                hit = &e;
        if (!hit) {
            if (!cap_stream_) SG_HIP(hipStreamCreateWithFlags(&cap_stream_, hipStreamNonBlocking));
            SG_HIP(hipStreamBeginCapture(cap_stream_, hipStreamCaptureModeThreadLocal));
            try {
                dispatch_(H, S, W, U, info, batch, cap_stream_, ws);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(cap_stream_, &g);
                if (g) (void)hipGraphDestroy(g);
                if (cur != device_) (void)hipSetDevice(cur);
                throw;
            }
            hipGraph_t g = nullptr;
            SG_HIP(hipStreamEndCapture(cap_stream_, &g));
            hipGraphExec_t ex = nullptr;
            const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            SG_HIP(ie);
            if (graph_cache_.size() >= 8) clear_graphs_();
            graph_cache_.push_back({H, Hd, hds, S, W, U, info, batch, ex});
            hit = &graph_cache_.back();
            ++graph_captures_;
        } else {
            ++graph_replays_;
        }
        SG_HIP(hipGraphLaunch(hit->exec, stream));
    } else {
        mark_(stream);
        dispatch_(H, S, W, U, info, batch, stream, ws);
    }
    if (multi) SG_HIP(hipEventRecord(done_, stream));
    if (cur != device_) SG_HIP(hipSetDevice(cur));
}

void Solver::launch_unordered(const void* H, const void* S, double* W, void* U, int* info, int batch,
                              hipStream_t stream) {
    if (batch < 0) throw std::invalid_argument("batch_size < 0");
    if (batch == 0) return;
    int cur = 0;
    SG_HIP(hipGetDevice(&cur));
    if (cur != device_) SG_HIP(hipSetDevice(device_));
    ev_used_ = 0;
    const bool multi = large_ || mode_ != Mode::Unified;
    const Workspace ws = multi ? own_workspace_((size_t)batch) : Workspace{};
    mark_(stream);
    dispatch_(H, S, W, U, info, batch, stream, ws);
    if (cur != device_) SG_HIP(hipSetDevice(cur));
}

void Solver::launch_chunked(const void* H, const void* S, double* W, void* U, int* info, int total,
                            int chunk, const std::vector<hipStream_t>& streams) {
    if (total < 0) throw std::invalid_argument("total_matrices < 0");
    if (total == 0) return;
    if (chunk <= 0) throw std::invalid_argument("chunk_size must be > 0");
    if (streams.empty()) throw std::invalid_argument("at least one stream is required");
    const size_t mat = (size_t)n_ * n_ * 16;
    int k = 0;
    for (int off = 0; off < total; off += chunk, ++k) {
        const int cnt = std::min(chunk, total - off);
        launch(static_cast<const char*>(H) + (size_t)off * mat, static_cast<const char*>(S) + (size_t)off * mat,
               W + (size_t)off * n_, static_cast<char*>(U) + (size_t)off * mat, info + off, cnt,
               streams[(size_t)k % streams.size()]);
    }
}

void Solver::dispatch_(const void* H, const void* S, double* W, void* U, int* info, int batch,
                       hipStream_t stream, const Workspace& ws) {
    if (ws.hdiag && ws.hfull) {
        // Paths that read H: write the diagonal out in full first.
        const double* hd = ws.hdiag;
        long long hds = ws.hdiag_stride;
        void* hf = ws.hfull;
        void* ae[] = {&hd, &hds, &hf, &batch};
        launch_(k_expand_, (unsigned)batch, 256, stream, ae);
        mark_(stream);
        H = hf;
    }
    if (large_) launch_large_(H, S, W, U, info, batch, stream, ws);
    else if (mode_ == Mode::Unified) launch_unified_(H, S, W, U, info, batch, stream);
    else if (mode_ == Mode::Hybrid) launch_hybrid_(H, S, W, U, info, batch, stream, ws);
    else launch_mixed_(H, S, W, U, info, batch, stream, ws);
}

Solver::Workspace Solver::own_workspace_(size_t batch) {
    Workspace ws;
    if (large_) {
        ensure_large_workspace_(batch);
        ws.staged = gbuf_;
        ws.staged_batch = gbuf_batch_;
        return ws;
    }
    if (mode_ == Mode::Hybrid) {
        ensure_workspace_(batch);
        ws.split = lbuf_;
    } else if (mode_ == Mode::Mixed) {
        ensure_mixed_workspace_(batch);
        ws.mixed = mbuf_;
    }
    if (mode_ != Mode::Unified && batch >= (size_t)ql_log_min_batch_) {
        ensure_qlog_(batch);
        ws.qlog = qlog_;
    }
    return ws;
}

size_t Solver::scratch_bytes_needed(Mode mode, int n, size_t batch) {
    if (n > kMaxLdsOrder) return large_matrix_bytes(n);
    const size_t nn = (size_t)n;
    if (mode == Mode::Hybrid) return hybrid_layout(nn, batch).total;
    if (mode == Mode::Mixed) return mixed_layout(nn, batch).total;
    return 0;
}

void Solver::launch_in_scratch(const void* H, const void* S, double* W, void* U, int* info, int batch,
                               hipStream_t stream, void* scratch, size_t scratch_bytes) {
    if (batch < 0) throw std::invalid_argument("batch_size < 0");
    if (batch == 0) return;
    const size_t need = scratch_bytes_needed(mode_, n_, (size_t)batch);
    if (need > scratch_bytes) {
        const size_t per = scratch_bytes_needed(mode_, n_, 1);
        throw std::invalid_argument("batch of " + std::to_string(batch) + " matrices of order " +
                                    std::to_string(n_) + " needs " + std::to_string(need) +
                                    " B of scratch, more than the " + std::to_string(scratch_bytes) +
                                    " B available (at most about " + std::to_string(per ? scratch_bytes / per : 0) +
                                    " matrices per call in this mode)");
    }
    int cur = 0;
    SG_HIP(hipGetDevice(&cur));
    if (cur != device_) SG_HIP(hipSetDevice(device_));
    Workspace ws;
    if (large_) {
        ws.staged = scratch;
        ws.staged_batch = scratch_bytes / large_matrix_bytes(n_);
    } else if (mode_ == Mode::Hybrid) {
        ws.split = scratch;
    } else if (mode_ == Mode::Mixed) {
        ws.mixed = scratch;
    }
    const bool multi = large_ || mode_ != Mode::Unified;
    ev_used_ = 0;
    if (multi) SG_HIP(hipStreamWaitEvent(stream, done_, 0));
    mark_(stream);
    dispatch_(H, S, W, U, info, batch, stream, ws);
    if (multi) SG_HIP(hipEventRecord(done_, stream));
    if (cur != device_) SG_HIP(hipSetDevice(cur));
}

void Solver::clear_graphs_() {
    for (auto& e : graph_cache_) (void)hipGraphExecDestroy(e.exec);
    graph_cache_.clear();
}

void Solver::set_graphs(bool on) {
    graphs_ = on;
    if (!on) clear_graphs_();
}

void Solver::launch_unified_(const void* H, const void* S, double* W, void* U, int* info, int batch,
                             hipStream_t stream) {
    void* args[] = {&H, &S, &W, &U, &info, &batch};
    const unsigned grid = (unsigned)((batch + bpb_ - 1) / bpb_);
    launch_(fused_, grid, (unsigned)block_dim_, stream, args);
    mark_(stream);
}

void Solver::ensure_workspace_(size_t batch) {
    if (batch <= lbuf_batch_) return;
    clear_graphs_();              // captured launches hold the old workspace
    if (lbuf_) {
        SG_HIP(hipDeviceSynchronize());
        SG_HIP(hipFree(lbuf_));
        lbuf_ = nullptr;
        lbuf_batch_ = 0;
    }
    // Per matrix: W = L^-1, packed L and M, the tridiagonal's off-diagonal and Z (hybrid_layout).
    SG_HIP(hipMalloc(&lbuf_, hybrid_layout((size_t)n_, batch).total));
    lbuf_batch_ = batch;
}

// Split mode (reference: hybrid_solver.cpp:295-393), FP64 throughout: the
// reduction (Cholesky, L and W = L^-1, M = W H W^H on the matrix cores, the
// tridiagonalisation of M), QL, then the back transform (V = Q Z on the
// matrix cores, U = L^-H V by block back substitution); each kernel at the
// occupancy its own footprint allows.
void Solver::launch_hybrid_(const void* H, const void* S, double* W, void* U, int* info, int batch,
                            hipStream_t stream, const Workspace& ws) {
    const size_t n = (size_t)n_, B = (size_t)batch;
    const HybridLayout hl = hybrid_layout(n, B);
    char* base = static_cast<char*>(ws.split);
    void* Wd = base + hl.wd;
    void* Lp = base + hl.lp;
    void* Mp = base + hl.mp;
    double* E = reinterpret_cast<double*>(base + hl.e);
    double* Zw = reinterpret_cast<double*>(base + hl.z);
    const unsigned grid_b = (unsigned)batch;
    const unsigned grid_w = (unsigned)((batch + bpb_ - 1) / bpb_);
    const unsigned bd = (unsigned)block_dim_;
    // One wave per 16-row (or 16-column) tile for the matrix-core kernels.
    const unsigned bd_tiles = (unsigned)(64 * ((n_ + 15) / 16));
    long long wstride = (long long)((w_packed_ ? n * (n + 1) / 2 : n * n) * 16);
    void* a1[] = {&S, &W, &U, &info, &Wd, &wstride, &Lp, &batch};
    launch_(k_mx_chol_, grid_b, bd_tiles, stream, a1);
    mark_(stream);
    if (ws.hdiag && !ws.hfull) {
        const double* hd = ws.hdiag;
        long long hds = ws.hdiag_stride;
        void* ad[] = {&hd, &hds, &Wd, &info, &Mp, &batch};
        launch_(k_mx_formm_diag_, grid_b, bd_tiles, stream, ad);
    } else {
        void* af[] = {&H, &Wd, &info, &Mp, &batch};
        launch_(k_mx_formm_, grid_b, bd_tiles, stream, af);
    }
    mark_(stream);
    // Tridiagonalisation: diagonal to W, off-diagonal to E, reflectors to U.
    void* a2[] = {&Mp, &W, &U, &info, &E, &batch};
    launch_(k_front64_, grid_b, bd, stream, a2);
    mark_(stream);
    char* qlog = static_cast<char*>(ws.qlog);
    long long qstride = (long long)ql_log_bytes(n_);
    int cap = ql_sweep_cap_;
    int all = (batch < ql_log_min_batch_ || !qlog) ? 1 : 0;
    last_ql_log_ = !all;
    if (!all) {
        void* ar[] = {&W, &E, &info, &qlog, &qstride, &batch, &cap};
        launch_(k_qlrec_, (unsigned)((batch + ql_rec_lanes_ - 1) / ql_rec_lanes_), 64, stream, ar);
        mark_(stream);
        void* a3[] = {&W, &E, &Zw, &info, &qlog, &qstride, &batch};
        launch_(k_heev_, grid_w, bd, stream, a3);
        mark_(stream);
    }
    void* aq[] = {&W, &E, &Zw, &info, &qlog, &qstride, &batch, &all};
    launch_(k_qlfull_, grid_w, bd, stream, aq);
    mark_(stream);
    void* a4[] = {&U, &Zw, &info, &batch};
    launch_(k_back64_, grid_b, bd_tiles, stream, a4);
    mark_(stream);
    void* a5[] = {&Lp, &U, &info, &batch};
    launch_(k_backtri_, grid_b, bd_tiles, stream, a5);
    mark_(stream);
}

void Solver::ensure_mixed_workspace_(size_t batch) {
    if (batch <= mbuf_batch_) return;
    clear_graphs_();
    if (mbuf_) {
        SG_HIP(hipDeviceSynchronize());
        SG_HIP(hipFree(mbuf_));
        mbuf_ = nullptr;
        mbuf_batch_ = 0;
    }
    // complex128 W = L^-1 (room for the full matrix) and L (packed),
    // complex64 stash / V, float eigenvalues, off-diagonals and Z, M (packed
    // complex128).
    SG_HIP(hipMalloc(&mbuf_, mixed_layout((size_t)n_, batch).total));
    mbuf_batch_ = batch;
}

// Mixed precision: FP64 Cholesky, L and L^-1, FP32 tridiagonalisation / QL /
// V = Q Z (the split kernels built in FP32), FP64 refinement of (M, V) in the
// Cholesky frame on the matrix cores, then U = L^-H V by back substitution.
void Solver::launch_mixed_(const void* H, const void* S, double* W, void* U, int* info, int batch,
                           hipStream_t stream, const Workspace& ws) {
    const size_t n = (size_t)n_, B = (size_t)batch;
    // Offsets for this batch (the buffer holds at least this many matrices).
    const MixedLayout ml = mixed_layout(n, B);
    char* base = static_cast<char*>(ws.mixed);
    void* Wd = base + ml.wd;
    void* Lp = base + ml.lp;
    void* V32 = base + ml.v32;
    void* W32 = base + ml.w32;
    void* E32 = base + ml.e32;
    void* Z32 = base + ml.z32;
    void* Mp = base + ml.mp;
    const unsigned grid_b = (unsigned)batch;
    const unsigned grid_w = (unsigned)((batch + bpb_ - 1) / bpb_);
    const unsigned bd = (unsigned)block_dim_;
    // FP64 Cholesky and W = L^-1 (its lower triangle packed by columns, or in
    // full with the upper zero): the staged kernels' panel factorisation, one
    // workgroup per matrix of one wave per 16-row tile (the kernel's size for
    // orders up to 64).
    long long wstride = (long long)((w_packed_ ? n * (n + 1) / 2 : n * n) * 16);
    void* a1[] = {&S, &W, &U, &info, &Wd, &wstride, &Lp, &batch};
    const unsigned bd_chol = (unsigned)(64 * ((n_ + 15) / 16));
    launch_(k_mx_chol_, grid_b, bd_chol, stream, a1);
    mark_(stream);
    // M = W H W^H once in FP64 on the matrix cores (one wave per 16-column block).
    const unsigned bd_ref = (unsigned)(64 * ((n_ + 15) / 16));
    if (ws.hdiag && !ws.hfull) {
        // Diagonal H: M = W diag(h) W^H straight from the diagonal.
        const double* hd = ws.hdiag;
        long long hds = ws.hdiag_stride;
        void* ad[] = {&hd, &hds, &Wd, &info, &Mp, &batch};
        launch_(k_mx_formm_diag_, grid_b, bd_ref, stream, ad);
    } else {
        void* af[] = {&H, &Wd, &info, &Mp, &batch};
        launch_(k_mx_formm_, grid_b, bd_ref, stream, af);
    }
    mark_(stream);
    void* a2[] = {&Mp, &W32, &V32, &info, &E32, &batch};
    launch_(k_mx_front_, grid_b, bd, stream, a2);
    mark_(stream);
    char* qlog = static_cast<char*>(ws.qlog);
    long long qstride = (long long)ql_log_bytes(n_);
    int cap = ql_sweep_cap_;
    int all = (batch < ql_log_min_batch_ || !qlog) ? 1 : 0;
    last_ql_log_ = !all;
    if (!all) {
        void* ar[] = {&W32, &E32, &info, &qlog, &qstride, &batch, &cap};
        launch_(k_qlrec_, (unsigned)((batch + ql_rec_lanes_ - 1) / ql_rec_lanes_), 64, stream, ar);
        mark_(stream);
        void* a3[] = {&W32, &E32, &Z32, &info, &qlog, &qstride, &batch};
        launch_(k_ql32_, grid_w, bd, stream, a3);
        mark_(stream);
    }
    void* aq[] = {&W32, &E32, &Z32, &info, &qlog, &qstride, &batch, &all};
    launch_(k_qlfull32_, grid_w, bd, stream, aq);
    mark_(stream);
    // V = Q Z: one workgroup per matrix, one wave per 16-column panel.
    void* a4[] = {&V32, &Z32, &info, &batch};
    launch_(k_back32_, grid_b, bd_ref, stream, a4);
    mark_(stream);
    int iters = refine_iters_;
    double delta = refine_delta_;
    double tol = refine_tol_;
    void* a5[] = {&Mp, &Lp, &W, &U, &info, &V32, &W32, &batch, &iters, &delta, &tol};
    // The refinement runs one wave per 16-column block of the matrix.
    launch_(k_refine_, grid_b, bd_ref, stream, a5);
    mark_(stream);
}

size_t Solver::large_matrix_bytes(int n_) {
    // The staged kernels' per-matrix workspace (geneig_large.hip: WS_* and the
    // QL rotation log Q_*); bind_module_ checks it against the module's own.
    const size_t n = (size_t)n_;
    auto al = [](size_t x, size_t a) { return (x + a - 1) / a * a; };
    const size_t wsl = al(88 * n * n + 24 * n, 256);           // L packed (complex128)
    const size_t wsq = al(wsl + 8 * n * (n + 1), 256);
    const size_t q_rank = 16 + 8 * (8 * n);
    const size_t q_rot = al(q_rank + 4 * n, 16);
    const size_t q_bytes = q_rot + 16 * (2 * n * n);
    return al(wsq + q_bytes, 256);
}

size_t Solver::ql_log_bytes(int n_) {
    // geneig_kernels.hip: QG_* (sweep windows, ranks, eigenvalues, rotations).
    const size_t n = (size_t)n_;
    auto al = [](size_t x, size_t a) { return (x + a - 1) / a * a; };
    const size_t rank = 16 + 8 * (8 * n);
    const size_t lam = al(rank + 4 * n, 16);
    const size_t rot = al(lam + 8 * n, 16);
    return al(rot + 16 * (2 * n * n), 256);
}

void Solver::ensure_hbuf_(size_t batch) {
    if (batch <= hbuf_batch_) return;
    clear_graphs_();
    if (hbuf_) {
        SG_HIP(hipDeviceSynchronize());
        SG_HIP(hipFree(hbuf_));
        hbuf_ = nullptr;
        hbuf_batch_ = 0;
    }
    SG_HIP(hipMalloc(&hbuf_, batch * (size_t)n_ * n_ * 16));
    hbuf_batch_ = batch;
}

void Solver::ensure_qlog_(size_t batch) {
    if (batch <= qlog_batch_) return;
    clear_graphs_();
    if (qlog_) {
        SG_HIP(hipDeviceSynchronize());
        SG_HIP(hipFree(qlog_));
        qlog_ = nullptr;
        qlog_batch_ = 0;
    }
    SG_HIP(hipMalloc(&qlog_, batch * ql_log_bytes(n_)));
    qlog_batch_ = batch;
}

size_t Solver::large_chunk(int n) { return std::max<size_t>(1, kLargeRoundBytes / large_matrix_bytes(n)); }

void Solver::ensure_large_workspace_(size_t batch) {
    batch = std::min(batch, large_chunk(n_));
    if (batch <= gbuf_batch_) return;
    clear_graphs_();
    if (gbuf_) {
        SG_HIP(hipDeviceSynchronize());
        SG_HIP(hipFree(gbuf_));
        gbuf_ = nullptr;
        gbuf_batch_ = 0;
    }
    SG_HIP(hipMalloc(&gbuf_, batch * large_matrix_bytes(n_)));
    gbuf_batch_ = batch;
}

// Orders above 64, every mode: the staged kernels, one workgroup per matrix,
// in rounds of at most large_chunk(n) matrices. FP64 modes: Cholesky and
// L^-1, M = W H W^H, tridiagonalisation, QL, V = Q Z and U = W^H V, all in
// FP64. Mixed: the middle three in FP32, then FP64 refinement of (M, V).
void Solver::launch_large_(const void* H, const void* S, double* W, void* U, int* info, int batch,
                           hipStream_t stream, const Workspace& ws_) {
    const size_t n = (size_t)n_, nn = n * n;
    const long long wss = (long long)large_matrix_bytes(n_);
    const unsigned bd = 256, bd_ql = (unsigned)(64 * ((n_ + 63) / 64));
    const int chunk = (int)std::min<size_t>(ws_.staged_batch, (size_t)batch);
    char* ws = static_cast<char*>(ws_.staged);
    for (int b0 = 0; b0 < batch; b0 += chunk) {
        int cnt = std::min(chunk, batch - b0);
        const void* Hc = static_cast<const char*>(H) + (size_t)b0 * nn * 16;
        const void* Sc = static_cast<const char*>(S) + (size_t)b0 * nn * 16;
        double* Wc = W + (size_t)b0 * n;
        void* Uc = static_cast<char*>(U) + (size_t)b0 * nn * 16;
        int* Ic = info + b0;
        const bool marks = (b0 == 0);
        auto mark = [&]() {
            if (marks) mark_(stream);
        };
        void* no_l = nullptr;     // the staged back-transform reads W
        void* a1[] = {&Sc, &Wc, &Uc, &Ic, &ws, (void*)&wss, &no_l, &cnt};
        launch_(k_lg_chol_, (unsigned)cnt, bd, stream, a1);
        mark();
        void* a2[] = {&Hc, &Ic, &ws, (void*)&wss, &cnt};
        launch_(k_lg_formm_, (unsigned)cnt, bd, stream, a2);
        mark();
        void* a3[] = {&Ic, &ws, (void*)&wss, &cnt};
        launch_(k_lg_hetrd_, (unsigned)cnt, lg_hetrd_threads_, stream, a3);
        mark();
        // QL eigenvalues: the caller's W (FP64), or the workspace's FP32 slot.
        void* wout = Wc;
        long long wstride = (long long)n;
        if (mode_ == Mode::Mixed) {
            wout = ws + 88 * nn + 8 * n;
            wstride = wss / 4;
        }
        int cap = ql_sweep_cap_;
        void* a4[] = {&wout, &wstride, &Ic, &ws, (void*)&wss, &cnt, &cap};
        if (k_lg_qll_ && cnt >= lanes_min_batch_)
            launch_(k_lg_qll_, (unsigned)((cnt + lg_qll_lanes_ - 1) / lg_qll_lanes_), 64, stream, a4);
        else
            launch_(k_lg_ql_, (unsigned)cnt, 64, stream, a4);
        mark();
        // Matrices whose rotation log overflowed (none, usually: each
        // workgroup returns after one header read), then every other one.
        launch_(k_lg_qlzi_, (unsigned)cnt, bd_ql, stream, a4);
        launch_(k_lg_qlz_, (unsigned)cnt, bd_ql, stream, a4);
        mark();
        launch_(k_lg_back_, (unsigned)cnt, lg_back_threads_, stream, a3);
        mark();
        if (mode_ == Mode::Mixed) {
            int iters = refine_iters_;
            double delta = refine_delta_;
            double tol = refine_tol_;
            void* a6[] = {&Wc, &Uc, &Ic, &ws, (void*)&wss, &cnt, &iters, &delta, &tol};
            launch_(k_lg_refine_, (unsigned)cnt, bd, stream, a6);
        } else {
            void* a6[] = {&Uc, &Ic, &ws, (void*)&wss, &cnt};
            launch_(k_lg_backtri_, (unsigned)cnt, bd, stream, a6);
        }
        mark();
    }
}


size_t Solver::workspace_bytes(Mode mode, int n_, size_t batch) {
    if (n_ > kMaxLdsOrder) return std::min(batch, large_chunk(n_)) * large_matrix_bytes(n_);
    const size_t n = (size_t)n_;
    // The QL rotation logs only for batches that run QL through them.
    const size_t qlog = batch >= (size_t)ql_log_min_batch_default(mode) ? batch * ql_log_bytes(n_) : 0;
    if (mode == Mode::Hybrid) return hybrid_layout(n, batch).total + qlog;
    if (mode == Mode::Mixed) return mixed_layout(n, batch).total + qlog;
    return 0;
}

size_t Solver::reserve(size_t batch) {
    int cur = 0;
    SG_HIP(hipGetDevice(&cur));
    if (cur != device_) SG_HIP(hipSetDevice(device_));
    if (large_) ensure_large_workspace_(batch);
    else if (mode_ == Mode::Hybrid) ensure_workspace_(batch);
    else if (mode_ == Mode::Mixed) ensure_mixed_workspace_(batch);
    if (!large_ && mode_ != Mode::Unified && batch >= (size_t)ql_log_min_batch_) ensure_qlog_(batch);
    if (cur != device_) SG_HIP(hipSetDevice(cur));
    return workspace_bytes(mode_, n_, batch);
}

void Solver::set_stage_timing(bool on) {
    timing_ = on;
    if (on && ev_.empty()) {
        ev_.resize(12);
        ev_fn_.assign(12, nullptr);
        for (auto& e : ev_) SG_HIP(hipEventCreate(&e));
    }
    ev_used_ = 0;
}

void Solver::launch_(hipFunction_t f, unsigned grid, unsigned block, hipStream_t stream, void** args) {
    SG_HIP(hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, 0, stream, args, nullptr));
    last_fn_ = f;
}

void Solver::mark_(hipStream_t stream) {
    if (!timing_ || ev_used_ >= (int)ev_.size()) return;
    // Event i closes the stage launched last before it (event 0 opens the launch).
    ev_fn_[ev_used_] = ev_used_ ? last_fn_ : nullptr;
    SG_HIP(hipEventRecord(ev_[ev_used_++], stream));
}

const char* Solver::stage_label(int i) const {
    if (i < 0 || i + 1 >= ev_used_) return "";
    hipFunction_t f = ev_fn_[i + 1];
    for (const auto* m : {module_.get(), module32_.get()}) {
        if (!m) continue;
        for (const auto& kv : m->functions)
            if (kv.second == f) return kv.first.c_str();
    }
    return "";
}

int Solver::stage_times(double* ms, int n) {
    if (!timing_ || ev_used_ < 2) return 0;
    SG_HIP(hipEventSynchronize(ev_[ev_used_ - 1]));
    const int stages = ev_used_ - 1;
    for (int i = 0; i < stages && i < n; ++i) {
        float t = 0.f;
        SG_HIP(hipEventElapsedTime(&t, ev_[i], ev_[i + 1]));
        ms[i] = t;
    }
    return stages;
}


}  // namespace sgeig
