// C ABI of libsunny_geneig.so (declared in include/sunny_geneig.h). Wraps the
// HIP solver in an opaque handle and turns C++ exceptions into status codes
// with a thread-local message, as the reference's src/common/c_api.cpp does.
#include "handles.hpp"

#include <exception>
#include <memory>
#include <new>
#include <string>
#include <vector>

namespace {

thread_local std::string g_last_error;
thread_local std::string g_cache_dir;

void set_error(const std::string& m) { g_last_error = m; }
void clear_error() { g_last_error.clear(); }

sgeig::CacheMode to_cache(sunny_geneig_cache_mode_t c) {
    switch (c) {
        case SUNNY_GENEIG_CACHE_NODISK: return sgeig::CacheMode::NoDisk;
        case SUNNY_GENEIG_CACHE_NOCACHE: return sgeig::CacheMode::NoCache;
        default: return sgeig::CacheMode::Auto;
    }
}

sunny_geneig_solver_t create_impl(int n, int device_id, int bpb, sunny_geneig_tuning_mode_t tuning,
                                  sunny_geneig_cache_mode_t cache, sunny_geneig_kernel_mode_t mode,
                                  const char* fn) {
    clear_error();
    try {
        auto h = std::make_unique<sunny_geneig_solver>();
        h->mode = mode;
        h->impl = std::make_unique<sgeig::Solver>(
            n, device_id, bpb,
            tuning == SUNNY_GENEIG_TUNING_AUTOTUNE ? sgeig::Tuning::Autotune : sgeig::Tuning::Suggested,
            to_cache(cache),
            mode == SUNNY_GENEIG_KERNEL_HYBRID  ? sgeig::Mode::Hybrid
            : mode == SUNNY_GENEIG_KERNEL_MIXED ? sgeig::Mode::Mixed
                                                : sgeig::Mode::Unified);
        // Per-device set-up of the no-allocation launch's scratch (once per
        // device, with that device current).
        sgeig::noalloc_setup(device_id);
        return h.release();
    } catch (const std::bad_alloc&) {
        set_error(std::string("out of memory in ") + fn);
    } catch (const std::exception& e) {
        set_error(std::string(fn) + ": " + e.what());
    } catch (...) {
        set_error(std::string("unknown C++ exception in ") + fn);
    }
    return nullptr;
}

}  // namespace

namespace sgeig {
void api_set_error(const std::string& m) { set_error(m); }
void api_clear_error() { clear_error(); }
}  // namespace sgeig

extern "C" {

sunny_geneig_solver_t sunny_geneig_create(int n, int device_id, int batches_per_block_request,
                                          sunny_geneig_tuning_mode_t tuning_mode,
                                          sunny_geneig_cache_mode_t cache_mode) {
    return create_impl(n, device_id, batches_per_block_request, tuning_mode, cache_mode,
                       SUNNY_GENEIG_KERNEL_UNIFIED, "sunny_geneig_create");
}

sunny_geneig_solver_t sunny_geneig_create_ex(int n, int device_id, int batches_per_block_request,
                                             sunny_geneig_tuning_mode_t tuning_mode,
                                             sunny_geneig_cache_mode_t cache_mode,
                                             sunny_geneig_kernel_mode_t kernel_mode) {
    if (kernel_mode != SUNNY_GENEIG_KERNEL_UNIFIED && kernel_mode != SUNNY_GENEIG_KERNEL_HYBRID &&
        kernel_mode != SUNNY_GENEIG_KERNEL_MIXED) {
        set_error("sunny_geneig_create_ex: unknown kernel_mode");
        return nullptr;
    }
    // As in the reference, the split mode ignores batches-per-block (so does
    // the mixed mode, built on the split kernels); the tuning mode applies to
    // every mode here.
    if (kernel_mode != SUNNY_GENEIG_KERNEL_UNIFIED)
        return create_impl(n, device_id, 0, tuning_mode, cache_mode, kernel_mode, "sunny_geneig_create_ex");
    return create_impl(n, device_id, batches_per_block_request, tuning_mode, cache_mode, kernel_mode,
                       "sunny_geneig_create_ex");
}

void sunny_geneig_destroy(sunny_geneig_solver_t solver) {
    try {
        delete solver;
    } catch (...) {
    }
}

sunny_geneig_status_t sunny_geneig_launch(sunny_geneig_solver_t solver, uintptr_t d_H, uintptr_t d_S,
                                          uintptr_t d_W, uintptr_t d_U, uintptr_t d_info,
                                          int batch_size, uintptr_t stream) {
    clear_error();
    if (!solver) {
        set_error("sunny_geneig_launch: solver is NULL");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    if (batch_size < 0) {
        set_error("sunny_geneig_launch: batch_size < 0");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    if (batch_size > 0 && (!d_H || !d_S || !d_W || !d_U || !d_info)) {
        set_error("sunny_geneig_launch: NULL device buffer");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    try {
        solver->impl->launch(reinterpret_cast<const void*>(d_H), reinterpret_cast<const void*>(d_S),
                             reinterpret_cast<double*>(d_W), reinterpret_cast<void*>(d_U),
                             reinterpret_cast<int*>(d_info), batch_size,
                             reinterpret_cast<hipStream_t>(stream));
        return SUNNY_GENEIG_OK;
    } catch (const std::exception& e) {
        set_error(std::string("sunny_geneig_launch: ") + e.what());
    } catch (...) {
        set_error("unknown C++ exception in sunny_geneig_launch");
    }
    return SUNNY_GENEIG_LAUNCH_ERROR;
}

sunny_geneig_status_t sunny_geneig_launch_hdiag(sunny_geneig_solver_t solver, uintptr_t d_Hdiag,
                                                int64_t hdiag_stride, uintptr_t d_S, uintptr_t d_W,
                                                uintptr_t d_U, uintptr_t d_info, int batch_size,
                                                uintptr_t stream) {
    clear_error();
    if (!solver) {
        set_error("sunny_geneig_launch_hdiag: solver is NULL");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    if (batch_size < 0 || hdiag_stride < 0) {
        set_error("sunny_geneig_launch_hdiag: batch_size and hdiag_stride must be >= 0");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    if (batch_size > 0 && (!d_Hdiag || !d_S || !d_W || !d_U || !d_info)) {
        set_error("sunny_geneig_launch_hdiag: NULL device buffer");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    try {
        solver->impl->launch_hdiag(reinterpret_cast<const double*>(d_Hdiag), (long long)hdiag_stride,
                                   reinterpret_cast<const void*>(d_S), reinterpret_cast<double*>(d_W),
                                   reinterpret_cast<void*>(d_U), reinterpret_cast<int*>(d_info), batch_size,
                                   reinterpret_cast<hipStream_t>(stream));
        return SUNNY_GENEIG_OK;
    } catch (const std::exception& e) {
        set_error(std::string("sunny_geneig_launch_hdiag: ") + e.what());
    } catch (...) {
        set_error("unknown C++ exception in sunny_geneig_launch_hdiag");
    }
    return SUNNY_GENEIG_LAUNCH_ERROR;
}

sunny_geneig_status_t sunny_geneig_launch_chunked(sunny_geneig_solver_t solver, uintptr_t d_H,
                                                  uintptr_t d_S, uintptr_t d_W, uintptr_t d_U,
                                                  uintptr_t d_info, int total_matrices, int chunk_size,
                                                  const uintptr_t* streams, int num_streams) {
    clear_error();
    if (!solver) {
        set_error("sunny_geneig_launch_chunked: solver is NULL");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    if (total_matrices < 0 || chunk_size <= 0 || num_streams <= 0 || !streams) {
        set_error("sunny_geneig_launch_chunked: need total_matrices >= 0, chunk_size > 0 and "
                  "num_streams > 0 streams");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    if (total_matrices > 0 && (!d_H || !d_S || !d_W || !d_U || !d_info)) {
        set_error("sunny_geneig_launch_chunked: NULL device buffer");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    try {
        std::vector<hipStream_t> st((size_t)num_streams);
        for (int i = 0; i < num_streams; ++i) st[(size_t)i] = reinterpret_cast<hipStream_t>(streams[i]);
        solver->impl->launch_chunked(reinterpret_cast<const void*>(d_H), reinterpret_cast<const void*>(d_S),
                                     reinterpret_cast<double*>(d_W), reinterpret_cast<void*>(d_U),
                                     reinterpret_cast<int*>(d_info), total_matrices, chunk_size, st);
        return SUNNY_GENEIG_OK;
    } catch (const std::exception& e) {
        set_error(std::string("sunny_geneig_launch_chunked: ") + e.what());
    } catch (...) {
        set_error("unknown C++ exception in sunny_geneig_launch_chunked");
    }
    return SUNNY_GENEIG_LAUNCH_ERROR;
}

int sunny_geneig_matrix_size(sunny_geneig_solver_t s) { return s ? s->impl->n() : -1; }
int sunny_geneig_batches_per_block(sunny_geneig_solver_t s) {
    return s ? s->impl->batches_per_block() : -1;
}
int sunny_geneig_block_dim_x(sunny_geneig_solver_t s) { return s ? s->impl->block_dim_x() : -1; }
unsigned int sunny_geneig_shared_mem_bytes(sunny_geneig_solver_t s) {
    return s ? s->impl->shared_mem_bytes() : 0u;
}

const char* sunny_geneig_last_error(void) { return g_last_error.c_str(); }

const char* sunny_geneig_path_name(sunny_geneig_solver_t s) {
    if (!s) return "";
    return s->mode == SUNNY_GENEIG_KERNEL_HYBRID  ? "hybrid"
           : s->mode == SUNNY_GENEIG_KERNEL_MIXED ? "mixed"
                                                  : "unified";
}

sunny_geneig_kernel_mode_t sunny_geneig_kernel_mode(sunny_geneig_solver_t s) {
    if (!s) return static_cast<sunny_geneig_kernel_mode_t>(-1);
    return s->mode;
}

int sunny_geneig_precompile(int n, int block_dim_x, const char* arch) {
    clear_error();
    try {
        if (n < 1 || n > sgeig::Solver::kMaxOrder) throw std::invalid_argument("n out of range");
        int bd = block_dim_x > 0 ? block_dim_x : sgeig::Solver::suggested_block_dim(n);
        const bool staged = block_dim_x < 0 && n <= sgeig::Solver::kMaxLdsOrder;
        if (n > sgeig::Solver::kMaxLdsOrder || staged) bd = 256;   // the staged kernels' only workgroup size
        sgeig::ModuleKey key{n, bd, arch && arch[0] ? arch : "gfx950"};
        key.staged = staged;
        sgeig::write_disk(key, sgeig::compile_code_object(key));
        key.fp32 = true;
        sgeig::write_disk(key, sgeig::compile_code_object(key));
        return SUNNY_GENEIG_OK;
    } catch (const sgeig::CompileError& e) {
        set_error(e.what());
        return SUNNY_GENEIG_COMPILE_ERROR;
    } catch (const std::exception& e) {
        set_error(e.what());
        return SUNNY_GENEIG_INVALID_ARG;
    }
}

int sunny_geneig_blocking(sunny_geneig_solver_t s) { return s ? (s->impl->staged() ? 1 : 0) : -1; }

const char* sunny_geneig_cache_layer(sunny_geneig_solver_t s) { return s ? s->impl->cache_layer() : ""; }

const char* sunny_geneig_cache_dir(void) {
    try {
        g_cache_dir = sgeig::cache_dir();
    } catch (...) {
        g_cache_dir.clear();
    }
    return g_cache_dir.c_str();
}

int sunny_geneig_set_stage_timing(sunny_geneig_solver_t s, int enable) {
    clear_error();
    if (!s) {
        set_error("sunny_geneig_set_stage_timing: solver is NULL");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    try {
        s->impl->set_stage_timing(enable != 0);
        return SUNNY_GENEIG_OK;
    } catch (const std::exception& e) {
        set_error(std::string("sunny_geneig_set_stage_timing: ") + e.what());
        return SUNNY_GENEIG_DEVICE_ERROR;
    }
}

int sunny_geneig_stage_times(sunny_geneig_solver_t s, double* ms, const char** names, int max_stages) {
    clear_error();
    if (!s) {
        set_error("sunny_geneig_stage_times: solver is NULL");
        return -1;
    }
    try {
        double tmp[12];
        const int n = s->impl->stage_times(tmp, 12);
        for (int i = 0; i < n && i < max_stages; ++i) {
            if (ms) ms[i] = tmp[i];
            if (names)
                names[i] = s->impl->stage_label(i);
        }
        return n;
    } catch (const std::exception& e) {
        set_error(std::string("sunny_geneig_stage_times: ") + e.what());
        return -1;
    }
}

int sunny_geneig_set_graphs(sunny_geneig_solver_t s, int enable) {
    clear_error();
    if (!s) {
        set_error("sunny_geneig_set_graphs: solver is NULL");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    s->impl->set_graphs(enable != 0);
    return SUNNY_GENEIG_OK;
}

int sunny_geneig_graph_stats(sunny_geneig_solver_t s, long long* captures, long long* replays) {
    clear_error();
    if (!s) {
        set_error("sunny_geneig_graph_stats: solver is NULL");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    if (captures) *captures = s->impl->graph_captures();
    if (replays) *replays = s->impl->graph_replays();
    return SUNNY_GENEIG_OK;
}

sunny_geneig_varlen_t sunny_geneig_varlen_create(int max_n, int device_id, const int* bucket_orders,
                                                 int num_buckets, sunny_geneig_cache_mode_t cache_mode) {
    return sunny_geneig_varlen_create_ex(max_n, device_id, bucket_orders, num_buckets, cache_mode,
                                         SUNNY_GENEIG_TUNING_SUGGESTED);
}

sunny_geneig_varlen_t sunny_geneig_varlen_create_ex(int max_n, int device_id, const int* bucket_orders,
                                                    int num_buckets, sunny_geneig_cache_mode_t cache_mode,
                                                    sunny_geneig_tuning_mode_t tuning_mode) {
    clear_error();
    try {
        if (num_buckets < 0) throw std::invalid_argument("num_buckets < 0");
        if (tuning_mode != SUNNY_GENEIG_TUNING_SUGGESTED && tuning_mode != SUNNY_GENEIG_TUNING_AUTOTUNE)
            throw std::invalid_argument("unknown tuning mode");
        std::vector<int> buckets;
        if (bucket_orders && num_buckets > 0) buckets.assign(bucket_orders, bucket_orders + num_buckets);
        auto h = std::make_unique<sunny_geneig_varlen_solver>();
        h->impl = std::make_unique<sgeig::VarlenSolver>(
            max_n, device_id, buckets, to_cache(cache_mode),
            tuning_mode == SUNNY_GENEIG_TUNING_AUTOTUNE ? sgeig::Tuning::Autotune : sgeig::Tuning::Suggested);
        return h.release();
    } catch (const std::bad_alloc&) {
        set_error("out of memory in sunny_geneig_varlen_create");
    } catch (const std::exception& e) {
        set_error(std::string("sunny_geneig_varlen_create: ") + e.what());
    } catch (...) {
        set_error("unknown C++ exception in sunny_geneig_varlen_create");
    }
    return nullptr;
}

void sunny_geneig_varlen_destroy(sunny_geneig_varlen_t solver) {
    try {
        delete solver;
    } catch (...) {
    }
}

sunny_geneig_status_t sunny_geneig_varlen_launch(sunny_geneig_varlen_t solver, uintptr_t d_H,
                                                 uintptr_t d_S, uintptr_t d_W, uintptr_t d_U,
                                                 uintptr_t d_info, const int* n_host, int count,
                                                 uintptr_t stream) {
    clear_error();
    if (!solver) {
        set_error("sunny_geneig_varlen_launch: solver is NULL");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    if (count < 0) {
        set_error("sunny_geneig_varlen_launch: count < 0");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    if (count > 0 && (!d_H || !d_S || !d_W || !d_U || !d_info || !n_host)) {
        set_error("sunny_geneig_varlen_launch: NULL buffer");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    try {
        solver->impl->launch(reinterpret_cast<const void*>(d_H), reinterpret_cast<const void*>(d_S),
                             reinterpret_cast<double*>(d_W), reinterpret_cast<void*>(d_U),
                             reinterpret_cast<int*>(d_info), n_host, count,
                             reinterpret_cast<hipStream_t>(stream));
        return SUNNY_GENEIG_OK;
    } catch (const std::invalid_argument& e) {
        set_error(std::string("sunny_geneig_varlen_launch: ") + e.what());
        return SUNNY_GENEIG_INVALID_ARG;
    } catch (const std::exception& e) {
        set_error(std::string("sunny_geneig_varlen_launch: ") + e.what());
    } catch (...) {
        set_error("unknown C++ exception in sunny_geneig_varlen_launch");
    }
    return SUNNY_GENEIG_LAUNCH_ERROR;
}

sunny_geneig_session_t sunny_geneig_session_create(int device_id, sunny_geneig_kernel_mode_t kernel_mode,
                                                   const int* orders, int num_orders, int max_count,
                                                   int num_slots, uint64_t hbm_budget_bytes,
                                                   sunny_geneig_cache_mode_t cache_mode) {
    clear_error();
    try {
        if (!orders || num_orders < 1) throw std::invalid_argument("no matrix orders to serve");
        if (kernel_mode != SUNNY_GENEIG_KERNEL_UNIFIED && kernel_mode != SUNNY_GENEIG_KERNEL_HYBRID &&
            kernel_mode != SUNNY_GENEIG_KERNEL_MIXED)
            throw std::invalid_argument("unknown kernel_mode");
        const sgeig::Mode mode = kernel_mode == SUNNY_GENEIG_KERNEL_HYBRID  ? sgeig::Mode::Hybrid
                                 : kernel_mode == SUNNY_GENEIG_KERNEL_MIXED ? sgeig::Mode::Mixed
                                                                            : sgeig::Mode::Unified;
        auto h = std::make_unique<sunny_geneig_session>();
        h->impl = std::make_unique<sgeig::ServingSession>(
            device_id, mode, std::vector<int>(orders, orders + num_orders), max_count, num_slots,
            (size_t)hbm_budget_bytes, to_cache(cache_mode));
        return h.release();
    } catch (const std::bad_alloc&) {
        set_error("out of memory in sunny_geneig_session_create");
    } catch (const std::exception& e) {
        set_error(std::string("sunny_geneig_session_create: ") + e.what());
    } catch (...) {
        set_error("unknown C++ exception in sunny_geneig_session_create");
    }
    return nullptr;
}

void sunny_geneig_session_destroy(sunny_geneig_session_t s) {
    try {
        delete s;
    } catch (...) {
    }
}

int64_t sunny_geneig_session_submit(sunny_geneig_session_t s, int n, int count, uintptr_t H, uintptr_t S,
                                    uintptr_t W, uintptr_t U, uintptr_t info, int on_device,
                                    uintptr_t stream) {
    clear_error();
    if (!s) {
        set_error("sunny_geneig_session_submit: session is NULL");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    try {
        return s->impl->submit(n, count, reinterpret_cast<const void*>(H), reinterpret_cast<const void*>(S),
                               reinterpret_cast<double*>(W), reinterpret_cast<void*>(U),
                               reinterpret_cast<int*>(info), on_device != 0,
                               reinterpret_cast<hipStream_t>(stream));
    } catch (const std::invalid_argument& e) {
        set_error(std::string("sunny_geneig_session_submit: ") + e.what());
        return SUNNY_GENEIG_INVALID_ARG;
    } catch (const std::exception& e) {
        set_error(std::string("sunny_geneig_session_submit: ") + e.what());
        return SUNNY_GENEIG_LAUNCH_ERROR;
    }
}

int sunny_geneig_session_set_policy(sunny_geneig_session_t s, int policy, int slice) {
    clear_error();
    if (!s) {
        set_error("sunny_geneig_session_set_policy: session is NULL");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    if (policy != 0 && policy != 1) {
        set_error("sunny_geneig_session_set_policy: policy must be 0 (arrival order) or 1 (least work left)");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    try {
        s->impl->set_policy(static_cast<sgeig::Policy>(policy), slice);
        return SUNNY_GENEIG_OK;
    } catch (const std::exception& e) {
        set_error(std::string("sunny_geneig_session_set_policy: ") + e.what());
        return SUNNY_GENEIG_INVALID_ARG;
    }
}

int sunny_geneig_session_poll(sunny_geneig_session_t s, int64_t id) {
    clear_error();
    if (!s) {
        set_error("sunny_geneig_session_poll: session is NULL");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    try {
        return s->impl->poll(id);
    } catch (const std::exception& e) {
        set_error(std::string("sunny_geneig_session_poll: ") + e.what());
        return SUNNY_GENEIG_INVALID_ARG;
    }
}

int sunny_geneig_session_wait(sunny_geneig_session_t s, int64_t id, double* latency_ms, double* service_ms) {
    clear_error();
    if (!s) {
        set_error("sunny_geneig_session_wait: session is NULL");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    try {
        s->impl->wait(id, latency_ms, service_ms);
        return SUNNY_GENEIG_OK;
    } catch (const std::invalid_argument& e) {
        set_error(std::string("sunny_geneig_session_wait: ") + e.what());
        return SUNNY_GENEIG_INVALID_ARG;
    } catch (const std::exception& e) {
        set_error(std::string("sunny_geneig_session_wait: ") + e.what());
        return SUNNY_GENEIG_LAUNCH_ERROR;
    }
}

int sunny_geneig_session_stats(sunny_geneig_session_t s, int64_t* submitted, int64_t* completed,
                               uint64_t* hbm_peak, int* max_concurrent) {
    clear_error();
    if (!s) {
        set_error("sunny_geneig_session_stats: session is NULL");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    const sgeig::ServingStats st = s->impl->stats();
    if (submitted) *submitted = st.submitted;
    if (completed) *completed = st.completed;
    if (hbm_peak) *hbm_peak = st.hbm_peak;
    if (max_concurrent) *max_concurrent = st.max_concurrent;
    return SUNNY_GENEIG_OK;
}

uint64_t sunny_geneig_session_request_bytes(int n, int count) {
    if (n < 1 || count < 1) return 0;
    return sgeig::ServingSession::request_bytes(n, count);
}

int sunny_geneig_varlen_set_graphs(sunny_geneig_varlen_t s, int enable) {
    clear_error();
    if (!s) {
        set_error("sunny_geneig_varlen_set_graphs: solver is NULL");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    try {
        s->impl->set_graphs(enable != 0);
        return SUNNY_GENEIG_OK;
    } catch (const std::exception& e) {
        set_error(std::string("sunny_geneig_varlen_set_graphs: ") + e.what());
        return SUNNY_GENEIG_DEVICE_ERROR;
    }
}

int sunny_geneig_varlen_graph_stats(sunny_geneig_varlen_t s, long long* captures, long long* replays) {
    clear_error();
    if (!s) {
        set_error("sunny_geneig_varlen_graph_stats: solver is NULL");
        return SUNNY_GENEIG_INVALID_ARG;
    }
    if (captures) *captures = s->impl->graph_captures();
    if (replays) *replays = s->impl->graph_replays();
    return SUNNY_GENEIG_OK;
}

int sunny_geneig_varlen_num_buckets(sunny_geneig_varlen_t s) {
    return s ? (int)s->impl->buckets().size() : -1;
}

int sunny_geneig_varlen_bucket_order(sunny_geneig_varlen_t s, int i) {
    if (!s || i < 0 || i >= (int)s->impl->buckets().size()) return -1;
    return s->impl->buckets()[i];
}

int sunny_geneig_varlen_bucket_waves(sunny_geneig_varlen_t s, int i) {
    if (!s || i < 0 || i >= (int)s->impl->buckets().size()) return -1;
    return s->impl->bucket_waves(i);
}

}  // extern "C"
