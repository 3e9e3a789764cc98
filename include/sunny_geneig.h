/* C ABI of libsunny_geneig.so: batched generalized Hermitian-definite
 * eigensolve H u = lambda S u on AMD Instinct MI355X (gfx950).
 *
 * Drop-in for the reference's src/common/c_api.h: the first eleven entry
 * points and the four enums keep the reference's names, signatures, values
 * and semantics (each declaration cites the line it replaces). Device
 * pointers and the stream cross the boundary as uintptr_t: a HIP device
 * pointer and a hipStream_t (0 = default stream).
 *
 * All functions are nothrow. A failure returns a status (or NULL) and leaves
 * a message in sunny_geneig_last_error() (thread-local).
 *
 * Buffers (caller-owned, column-major, element (i,j) of matrix b at
 * b*n*n + j*n + i):
 *   d_H, d_S : complex128 [batch][n][n], Hermitian; read only.
 *   d_W      : float64    [batch][n], eigenvalues ascending.
 *   d_U      : complex128 [batch][n][n], eigenvectors, U^H S U = I.
 *   d_info   : int32      [batch]; 0 = success, n + k = leading minor k of S
 *              not positive definite (zhegv INFO), 1..n-1 = QL did not
 *              converge (that many off-diagonals left), n = (MIXED mode only)
 *              the FP64 refinement's last correction was still above its
 *              tolerance (W and U hold the last iterate).
 */
#ifndef SUNNY_GENEIG_H
#define SUNNY_GENEIG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sunny_geneig_solver* sunny_geneig_solver_t;

/* Reference: src/common/c_api.h:24-30. */
typedef enum {
    SUNNY_GENEIG_OK            =  0,
    SUNNY_GENEIG_INVALID_ARG   = -1,
    SUNNY_GENEIG_DEVICE_ERROR  = -2,
    SUNNY_GENEIG_COMPILE_ERROR = -3,
    SUNNY_GENEIG_LAUNCH_ERROR  = -4
} sunny_geneig_status_t;

/* Reference: src/common/c_api.h:32-36. AUTO = in-process + on-disk code-object
 * cache; NODISK = in-process only; NOCACHE = compile every time. */
typedef enum {
    SUNNY_GENEIG_CACHE_AUTO    = 0,
    SUNNY_GENEIG_CACHE_NODISK  = 1,
    SUNNY_GENEIG_CACHE_NOCACHE = 2
} sunny_geneig_cache_mode_t;

/* Reference: src/common/c_api.h:38-41. SUGGESTED = built-in workgroup size
 * model; AUTOTUNE = timed, correctness-gated sweep (winner persisted). */
typedef enum {
    SUNNY_GENEIG_TUNING_SUGGESTED = 0,
    SUNNY_GENEIG_TUNING_AUTOTUNE  = 1
} sunny_geneig_tuning_mode_t;

/* Reference: src/common/c_api.h:43-46. UNIFIED = one fused kernel per batch;
 * HYBRID = split mode, FP64: reduction kernels (Cholesky and W = L^-1,
 * M = W H W^H, tridiagonalisation), batched QL kernels, batched
 * back-transform kernels (V = Q Z, U = W^H V), all written for gfx950 by this
 * project. MIXED is additive (no counterpart in the reference): the same FP64
 * Cholesky and M, FP32 tridiagonalisation / QL / V = Q Z, then FP64
 * eigenpair refinement; same buffers and outputs as the other modes. */
typedef enum {
    SUNNY_GENEIG_KERNEL_UNIFIED = 0,
    SUNNY_GENEIG_KERNEL_HYBRID  = 1,
    SUNNY_GENEIG_KERNEL_MIXED   = 2
} sunny_geneig_kernel_mode_t;

/* Reference: src/common/c_api.h:50-55 (c_api.cpp:76-107). Returns NULL on
 * failure. batches_per_block_request <= 0 selects automatically; 1..4 is
 * honoured (one wave per matrix); a larger request returns NULL with a message
 * naming the limit, as the reference does for an explicit BPB it cannot build
 * (unified_solver.cpp:1021-1028; its ladder goes to 16, this kernel's to 4). */
sunny_geneig_solver_t sunny_geneig_create(
    int                         n,
    int                         device_id,
    int                         batches_per_block_request,
    sunny_geneig_tuning_mode_t  tuning_mode,
    sunny_geneig_cache_mode_t   cache_mode);

/* Reference: src/common/c_api.h:57-69 (c_api.cpp:109-137). Unlike the
 * reference, AUTOTUNE also applies to the HYBRID and MIXED modes here: the
 * timed, correctness-gated sweep covers the workgroup size (= matrices per
 * workgroup) and, from order 48, LDS-resident versus HBM-staged blocking. */
sunny_geneig_solver_t sunny_geneig_create_ex(
    int                         n,
    int                         device_id,
    int                         batches_per_block_request,
    sunny_geneig_tuning_mode_t  tuning_mode,
    sunny_geneig_cache_mode_t   cache_mode,
    sunny_geneig_kernel_mode_t  kernel_mode);

/* Reference: src/common/c_api.h:71-72 (c_api.cpp:139-145). NULL-safe. */
void sunny_geneig_destroy(sunny_geneig_solver_t solver);

/* Reference: src/common/c_api.h:74-85 (c_api.cpp:147-186). Asynchronous on
 * `stream`. NULL solver or batch_size < 0 -> INVALID_ARG; batch_size == 0 is a
 * no-op. */
sunny_geneig_status_t sunny_geneig_launch(
    sunny_geneig_solver_t solver,
    uintptr_t             d_H,
    uintptr_t             d_S,
    uintptr_t             d_W,
    uintptr_t             d_U,
    uintptr_t             d_info,
    int                   batch_size,
    uintptr_t             stream);

/* Reference: src/common/c_api.h:87-91 (c_api.cpp:188-206). -1 / 0 on NULL.
 * batches_per_block is 1 in the split (HYBRID) and MIXED modes, as the
 * reference's split mode reports (c_api.cpp:41-49); shared_mem_bytes is the
 * fused kernel's LDS, the largest of the split mode's reduction kernels', or
 * the largest of the mixed mode's kernels'. */
int          sunny_geneig_matrix_size      (sunny_geneig_solver_t solver);
int          sunny_geneig_batches_per_block(sunny_geneig_solver_t solver);
int          sunny_geneig_block_dim_x      (sunny_geneig_solver_t solver);
unsigned int sunny_geneig_shared_mem_bytes (sunny_geneig_solver_t solver);

/* Reference: src/common/c_api.h:93-95 (c_api.cpp:208-210). Never NULL. */
const char*  sunny_geneig_last_error(void);

/* Reference: src/common/c_api.h:97-101 (c_api.cpp:212-220). "unified",
 * "hybrid", "mixed", or "" for NULL; kernel_mode returns -1 for NULL. */
const char*                sunny_geneig_path_name  (sunny_geneig_solver_t solver);
sunny_geneig_kernel_mode_t sunny_geneig_kernel_mode(sunny_geneig_solver_t solver);

/* ---- additive entry points (no counterpart in the reference ABI) --------- */

/* Chunked launch over several streams: contiguous chunks of chunk_size
 * matrices (the last one shorter) go round-robin to streams[0..num_streams),
 * each as one sunny_geneig_launch on its stream (pointers offset by the
 * chunk's first matrix). C-ABI form of UnifiedGeneigSolver::launch_chunked
 * (reference: src/unified/unified_solver.cpp:1078-1107) and of the offset
 * arithmetic of run_kernel_chunked! (SunnyGeneigBench.jl:191-251).
 * Asynchronous; the caller synchronises the streams. In the fused mode the
 * chunks may run concurrently; the split and mixed modes and the staged
 * orders share one workspace per handle, so their chunks run in submission
 * order whatever the stream. total_matrices == 0 is a no-op; a negative
 * total, chunk_size <= 0, num_streams <= 0 or a NULL streams array ->
 * INVALID_ARG. A stream value of 0 is the default stream. */
sunny_geneig_status_t sunny_geneig_launch_chunked(
    sunny_geneig_solver_t solver,
    uintptr_t             d_H,
    uintptr_t             d_S,
    uintptr_t             d_W,
    uintptr_t             d_U,
    uintptr_t             d_info,
    int                   total_matrices,
    int                   chunk_size,
    const uintptr_t*      streams,
    int                   num_streams);

/* Launch for a real diagonal H (no counterpart in the reference ABI): the
 * Sunny call pattern passes H = Sigma = diag(I, -I), the paraunitary metric,
 * with S = the dynamical matrix (julia/SunnyGeneigBench/scripts/
 * setup_sw08.jl:6-18). d_Hdiag is float64: matrix b's diagonal is the n
 * values at d_Hdiag + b * hdiag_stride (hdiag_stride 0 = one diagonal for
 * the whole batch). Results and errors as sunny_geneig_launch on the full H.
 * In the MIXED mode at orders up to 64, M = L^-1 H L^-H is formed from the
 * diagonal directly (one signed Gram product, H never materialised); the
 * other paths write the full H into a handle workspace first.
 * hdiag_stride < 0 or a NULL diagonal with batch_size > 0 -> INVALID_ARG. */
sunny_geneig_status_t sunny_geneig_launch_hdiag(
    sunny_geneig_solver_t solver,
    uintptr_t             d_Hdiag,
    int64_t               hdiag_stride,
    uintptr_t             d_S,
    uintptr_t             d_W,
    uintptr_t             d_U,
    uintptr_t             d_info,
    int                   batch_size,
    uintptr_t             stream);

/* No-allocation launch. No counterpart in the reference ABI: it is for
 * callers that cannot allocate device memory (code under hipGraph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Same arguments,
 * results and error behaviour as sunny_geneig_launch, but the split, mixed
 * and staged-order workspace is carved from a device array of
 * SUNNY_GENEIG_NOALLOC_SCRATCH_BYTES that the library holds on each GPU it
 * runs on (set up when a handle is created there) until the process ends; the
 * call allocates nothing and captures no graph. A batch whose workspace does
 * not fit returns INVALID_ARG with the limit in the message (at N = 54: about
 * 1,030 matrices per call in the mixed mode, 1,130 in the split mode; the
 * fused mode needs no workspace; orders above 64 run in rounds that fit).
 * Calls on one GPU share its array and are serialised on the device through
 * an event, whatever stream or handle they use. */
#define SUNNY_GENEIG_NOALLOC_SCRATCH_BYTES ((uint64_t)128 << 20)
sunny_geneig_status_t sunny_geneig_launch_noalloc(
    sunny_geneig_solver_t solver,
    uintptr_t             d_H,
    uintptr_t             d_S,
    uintptr_t             d_W,
    uintptr_t             d_U,
    uintptr_t             d_info,
    int                   batch_size,
    uintptr_t             stream);

/* Compile (hipRTC) and store in the on-disk cache the code objects for order n
 * with `block_dim_x` threads (0 = suggested; -1 = the HBM-staged pipeline at
 * an order up to 64, the autotuner's other blocking) for `arch` (e.g. "gfx950"): the
 * FP64 module and the FP32 module of the mixed mode, without touching a device. Returns 0 or a status. Lets a build step warm the
 * cache so the first create on a fresh machine skips compilation. */
int sunny_geneig_precompile(int n, int block_dim_x, const char* arch);

/* Blocking of the handle's kernels: 0 = LDS-resident (one matrix, or a few,
 * per workgroup with the pencil in LDS; orders up to 64), 1 = HBM-staged (a
 * kernel per stage handing over through a per-matrix workspace; every order
 * above 64, and at or below 64 when TUNING_AUTOTUNE found it faster or the
 * environment variable SUNNY_GENEIG_STAGED was set to 1 at creation). -1 for
 * NULL. No counterpart in the reference ABI. */
int sunny_geneig_blocking(sunny_geneig_solver_t solver);

/* Where the cache layer that served construction came from: "in-process",
 * "disk" or "compile" (reference: unified_solver.hpp:129-136). "" for NULL. */
const char* sunny_geneig_cache_layer(sunny_geneig_solver_t solver);

/* Resolved on-disk cache directory (reference: unified_solver.cpp:403-420). */
const char* sunny_geneig_cache_dir(void);

/* Stage timing (no counterpart in the reference ABI; its timing is done by
 * callers around whole launches, unified_solver.cpp:838-856). When enabled,
 * HIP events bracket every kernel a launch issues, on the launch's stream.
 * sunny_geneig_stage_times waits for the last launch, writes up to
 * max_stages durations (ms) and, if names is not NULL, the kernel names, and
 * returns the stage count (UNIFIED 1, HYBRID 3, MIXED 5) or 0 when there is
 * no timed launch; -1 on NULL. Not for concurrent launches of one handle. */
int sunny_geneig_set_stage_timing(sunny_geneig_solver_t solver, int enable);
int sunny_geneig_stage_times(sunny_geneig_solver_t solver, double* ms, const char** names,
                             int max_stages);

/* hipGraph replay (no counterpart in the reference ABI). The split, mixed and
 * staged launches issue several kernels; a handle captures them once per
 * (d_H, d_S, d_W, d_U, d_info, batch_size) into a hipGraph and replays it on
 * the caller's stream when those come back (up to 8 cached; none while stage
 * timing is on; the environment variable SUNNY_GENEIG_GRAPH set to 0 at
 * creation turns it off).
 * set_graphs(h, 0) launches the kernels directly; graph_stats reports
 * captures and replays so far. */
int sunny_geneig_set_graphs(sunny_geneig_solver_t solver, int enable);
int sunny_geneig_graph_stats(sunny_geneig_solver_t solver, long long* captures, long long* replays);

/* ---- uneven batches (no counterpart in the reference ABI) ----------------
 * One call solves `count` pencils of different orders n[k] in [1, max_n].
 * Every matrix goes to the smallest bucket order >= n[k] and is embedded in
 * that order as (H (+) pad I, S (+) I), which leaves its eigenpairs
 * unchanged. Every bucket runs the FP64 split mode of sunny_geneig_launch at
 * its order (the staged kernels above 64) on a padded copy, gathered and
 * scattered on the device in rounds (4096 matrices above 64; below, as many
 * as keep the padded H, S and U within 4 GiB, at most 65536).
 *
 * Packed buffers, matrix k column-major at element offset sum_{j<k} n[j]^2:
 *   d_H, d_S, d_U : complex128, sum n[k]^2 elements;
 *   d_W           : float64, sum n[k] elements (matrix k at sum_{j<k} n[j]);
 *   d_info        : int32 [count], same codes as sunny_geneig_launch.
 * bucket_orders: strictly ascending orders in [1, 192], the last >= max_n; NULL
 * (or num_buckets == 0) = multiples of 4 below min(max_n, 64), of 8 below
 * min(max_n, 128), of 16 below max_n, then max_n. */
typedef struct sunny_geneig_varlen_solver* sunny_geneig_varlen_t;

sunny_geneig_varlen_t sunny_geneig_varlen_create(int max_n, int device_id,
                                                 const int* bucket_orders, int num_buckets,
                                                 sunny_geneig_cache_mode_t cache_mode);

/* As sunny_geneig_varlen_create, with a tuning mode: AUTOTUNE gives every
 * bucket what the timed sweep of sunny_geneig_create_ex's AUTOTUNE picks for
 * the split mode at that order (workgroup size, and from order 48
 * LDS-resident or HBM-staged blocking; run once per order and architecture,
 * then read from the cache's table); SUGGESTED is sunny_geneig_varlen_create.
 * Buckets above 64 have one size. Counterpart: the reference tunes one order
 * per handle only (unified_solver.cpp:649-925). */
sunny_geneig_varlen_t sunny_geneig_varlen_create_ex(int max_n, int device_id,
                                                    const int* bucket_orders, int num_buckets,
                                                    sunny_geneig_cache_mode_t cache_mode,
                                                    sunny_geneig_tuning_mode_t tuning_mode);

void sunny_geneig_varlen_destroy(sunny_geneig_varlen_t solver);

/* Asynchronous on `stream`; n_host (host memory, `count` orders) is read
 * before return. Calls on one handle are ordered after each other whatever
 * the stream; the first call with given buffers and orders is captured into
 * one hipGraph (every bucket's gather, solve and scatter rounds) that later calls
 * with the same buffers and orders replay without any host-side sorting or
 * upload. A handle is not reentrant: one call at a time per handle.
 * count == 0 is a no-op; an order outside [1, max_n] -> INVALID_ARG. */
sunny_geneig_status_t sunny_geneig_varlen_launch(
    sunny_geneig_varlen_t solver,
    uintptr_t             d_H,
    uintptr_t             d_S,
    uintptr_t             d_W,
    uintptr_t             d_U,
    uintptr_t             d_info,
    const int*            n_host,
    int                   count,
    uintptr_t             stream);

/* hipGraph replay of uneven-batch calls (see sunny_geneig_varlen_launch):
 * enable = 0 launches the grids directly (as the environment variable
 * SUNNY_GENEIG_GRAPH set to 0 at creation does); graph_stats reports captures
 * and replays so far. Return 0 or INVALID_ARG. */
int sunny_geneig_varlen_set_graphs(sunny_geneig_varlen_t solver, int enable);
int sunny_geneig_varlen_graph_stats(sunny_geneig_varlen_t solver, long long* captures, long long* replays);

/* Number of buckets, the order of bucket i and its matrices per workgroup
 * in the QL grids (1 on the staged kernels; -1 on NULL / out of range). */
int sunny_geneig_varlen_num_buckets (sunny_geneig_varlen_t solver);
int sunny_geneig_varlen_bucket_order(sunny_geneig_varlen_t solver, int i);
int sunny_geneig_varlen_bucket_waves(sunny_geneig_varlen_t solver, int i);

/* ---- serving session (no counterpart in the reference ABI) ---------------
 * A request loop over resident solvers: requests (count pencils of one order
 * from `orders`) are queued on submission. The session cuts them into slices
 * (see sunny_geneig_session_set_policy) and starts a slice whenever one of
 * `num_slots` slots (a stream with its own solver handles, workspaces
 * reserved for max_count pencils at creation) is free and, for a request's
 * first slice, its device footprint fits `hbm_budget_bytes`. Slices run
 * concurrently: copy in, solve, copy out. With on_device != 0 the five
 * buffers are caller-owned device memory (no copies, no footprint); otherwise
 * they are host memory (pinned memory gives asynchronous copies), staged
 * through the session's device arena. Buffers must stay valid until the
 * request is done. Per-pencil results equal those of sunny_geneig_launch. */
typedef struct sunny_geneig_session* sunny_geneig_session_t;

sunny_geneig_session_t sunny_geneig_session_create(int device_id, sunny_geneig_kernel_mode_t kernel_mode,
                                                   const int* orders, int num_orders, int max_count,
                                                   int num_slots, uint64_t hbm_budget_bytes,
                                                   sunny_geneig_cache_mode_t cache_mode);

/* Drains queued and running requests, then frees everything. NULL-safe. */
void sunny_geneig_session_destroy(sunny_geneig_session_t session);

/* Returns the request id (>= 0) or a negative status (INVALID_ARG for an
 * order the session does not serve, a count outside [1, max_count], NULL
 * buffers, or a footprint larger than the budget). `stream` is the stream on
 * which the caller produced the buffers (a hipStream_t; 0 = the null stream):
 * the session's copies and kernels for the request wait for the work that
 * stream holds at submission. For host buffers filled synchronously pass 0
 * (no wait is recorded). */
int64_t sunny_geneig_session_submit(sunny_geneig_session_t session, int n, int count,
                                    uintptr_t H, uintptr_t S, uintptr_t W, uintptr_t U,
                                    uintptr_t info, int on_device, uintptr_t stream);

/* Scheduling: policy 0 serves requests in arrival order, 1 serves first the
 * request with the least work left (pencils not yet started x n^3, oldest
 * first on ties); slices of at most `slice` pencils (0 = whole requests) let
 * one request use several slots at once and let the scheduler switch
 * requests between slices. Default: policy 0, slice 0. Applies to slices
 * started after the call. Returns 0 or INVALID_ARG. */
int sunny_geneig_session_set_policy(sunny_geneig_session_t session, int policy, int slice);

/* 1 when the request is done, 0 while queued or running, a negative status
 * for an unknown id. */
int sunny_geneig_session_poll(sunny_geneig_session_t session, int64_t request_id);

/* Blocks until the request is done and forgets it; latency_ms = submission
 * to completion, service_ms = admission to completion (either may be NULL).
 * Returns 0 or a negative status (LAUNCH_ERROR if the request failed). */
int sunny_geneig_session_wait(sunny_geneig_session_t session, int64_t request_id,
                              double* latency_ms, double* service_ms);

/* Counters since creation: requests submitted / completed, the arena's
 * high-water mark in bytes and the most slices running at once. */
int sunny_geneig_session_stats(sunny_geneig_session_t session, int64_t* submitted, int64_t* completed,
                               uint64_t* hbm_peak_bytes, int* max_concurrent);

/* Device footprint of a host-buffer request: H, S, U, W and info. */
uint64_t sunny_geneig_session_request_bytes(int n, int count);

#ifdef __cplusplus
}
#endif

#endif /* SUNNY_GENEIG_H */
