"""Python host side of libsunny_geneig.so (MI355X / gfx950).

Mirrors the reference's Julia wrapper SunnyGeneigBench
(julia/SunnyGeneigBench/src/SunnyGeneigBench.jl): a `Solver(n; device_id,
bpb_request, tuning, cache, kernel)` object over the C ABI, `run_kernel`
(run_kernel! at :104-171) and `run_kernel_chunked` (:191-251). Device memory
and streams are torch tensors / streams; every call goes through the C ABI
into the HIP kernels. There is no CPU fallback: if the library is missing the
import of this package raises.

Array convention: a batch of B matrices of order n is a complex128 tensor of
shape (B, n, n) holding each matrix column-major, i.e. T[b, j, i] = A_b(i, j)
(the memory layout of the reference's (n, n, B) Julia arrays). W is float64
(B, n); info is int32 (B,).
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SUNNY_GENEIG_LIBPATH", os.path.join(_PKG, "lib", "libsunny_geneig.so"))
DEFAULT_CACHE_DIR = os.path.join(_PKG, "_cocache")

# Status / mode constants: values are ABI (include/sunny_geneig.h).
STATUS_OK = 0
STATUS_INVALID_ARG = -1
STATUS_DEVICE_ERROR = -2
STATUS_COMPILE_ERROR = -3
STATUS_LAUNCH_ERROR = -4
TUNING_SUGGESTED = 0
TUNING_AUTOTUNE = 1
CACHE_AUTO = 0
CACHE_NODISK = 1
CACHE_NOCACHE = 2
KERNEL_UNIFIED = 0
KERNEL_HYBRID = 1
KERNEL_MIXED = 2

EXPORTED_SYMBOLS = (
    "sunny_geneig_create", "sunny_geneig_create_ex", "sunny_geneig_destroy",
    "sunny_geneig_launch", "sunny_geneig_matrix_size", "sunny_geneig_batches_per_block",
    "sunny_geneig_block_dim_x", "sunny_geneig_shared_mem_bytes", "sunny_geneig_last_error",
    "sunny_geneig_path_name", "sunny_geneig_kernel_mode", "sunny_geneig_precompile",
    "sunny_geneig_cache_layer", "sunny_geneig_cache_dir",
    "sunny_geneig_varlen_create", "sunny_geneig_varlen_destroy", "sunny_geneig_varlen_launch",
    "sunny_geneig_varlen_num_buckets", "sunny_geneig_varlen_bucket_order",
    "sunny_geneig_varlen_create_ex", "sunny_geneig_varlen_bucket_waves",
    "sunny_geneig_set_stage_timing", "sunny_geneig_stage_times",
    "sunny_geneig_set_graphs", "sunny_geneig_graph_stats",
    "sunny_geneig_session_create", "sunny_geneig_session_destroy", "sunny_geneig_session_submit",
    "sunny_geneig_session_poll", "sunny_geneig_session_wait", "sunny_geneig_session_stats",
    "sunny_geneig_session_request_bytes", "sunny_geneig_launch_chunked",
    "sunny_geneig_session_set_policy", "sunny_geneig_launch_noalloc", "sunny_geneig_launch_hdiag",
    "sunny_geneig_blocking", "sunny_geneig_varlen_set_graphs", "sunny_geneig_varlen_graph_stats",
)

_lib = None


def lib() -> ctypes.CDLL:
    """Load libsunny_geneig.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libsunny_geneig.so not found at {LIB_PATH}; run "
                          "`python -m sunny_gen_eigensolver_amd.build` first")
    # The in-tree code-object cache ships warm; use it unless told otherwise.
    os.environ.setdefault("SUNNY_GENEIG_CACHE_DIR", DEFAULT_CACHE_DIR)
    # torch ships its own HIP runtime and finds it by the unversioned name, this
    # library by the versioned SONAME. Loaded first, torch's copy also serves
    # this library (same SONAME); loaded second, it would be a second runtime
    # in the process and torch's device initialisation fails.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, ci, cu = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64
    L.sunny_geneig_create.argtypes = [ci, ci, ci, ci, ci]
    L.sunny_geneig_create.restype = vp
    L.sunny_geneig_create_ex.argtypes = [ci, ci, ci, ci, ci, ci]
    L.sunny_geneig_create_ex.restype = vp
    L.sunny_geneig_destroy.argtypes = [vp]
    L.sunny_geneig_destroy.restype = None
    L.sunny_geneig_launch.argtypes = [vp, cu, cu, cu, cu, cu, ci, cu]
    L.sunny_geneig_launch.restype = ci
    L.sunny_geneig_launch_hdiag.argtypes = [vp, cu, ctypes.c_int64, cu, cu, cu, cu, ci, cu]
    L.sunny_geneig_launch_hdiag.restype = ci
    L.sunny_geneig_launch_noalloc.argtypes = [vp, cu, cu, cu, cu, cu, ci, cu]
    L.sunny_geneig_launch_noalloc.restype = ci
    L.sunny_geneig_launch_chunked.argtypes = [vp, cu, cu, cu, cu, cu, ci, ci, ctypes.POINTER(cu), ci]
    L.sunny_geneig_launch_chunked.restype = ci
    for name in ("sunny_geneig_matrix_size", "sunny_geneig_batches_per_block",
                 "sunny_geneig_block_dim_x", "sunny_geneig_kernel_mode", "sunny_geneig_blocking"):
        getattr(L, name).argtypes = [vp]
        getattr(L, name).restype = ci
    L.sunny_geneig_shared_mem_bytes.argtypes = [vp]
    L.sunny_geneig_shared_mem_bytes.restype = ctypes.c_uint
    L.sunny_geneig_last_error.argtypes = []
    L.sunny_geneig_last_error.restype = ctypes.c_char_p
    L.sunny_geneig_path_name.argtypes = [vp]
    L.sunny_geneig_path_name.restype = ctypes.c_char_p
    L.sunny_geneig_precompile.argtypes = [ci, ci, ctypes.c_char_p]
    L.sunny_geneig_precompile.restype = ci
    L.sunny_geneig_cache_layer.argtypes = [vp]
    L.sunny_geneig_cache_layer.restype = ctypes.c_char_p
    L.sunny_geneig_cache_dir.argtypes = []
    L.sunny_geneig_cache_dir.restype = ctypes.c_char_p
    pi = ctypes.POINTER(ctypes.c_int)
    L.sunny_geneig_varlen_create.argtypes = [ci, ci, pi, ci, ci]
    L.sunny_geneig_varlen_create.restype = vp
    L.sunny_geneig_varlen_create_ex.argtypes = [ci, ci, pi, ci, ci, ci]
    L.sunny_geneig_varlen_create_ex.restype = vp
    L.sunny_geneig_varlen_bucket_waves.argtypes = [vp, ci]
    L.sunny_geneig_varlen_bucket_waves.restype = ci
    L.sunny_geneig_varlen_destroy.argtypes = [vp]
    L.sunny_geneig_varlen_destroy.restype = None
    L.sunny_geneig_varlen_launch.argtypes = [vp, cu, cu, cu, cu, cu, pi, ci, cu]
    L.sunny_geneig_varlen_launch.restype = ci
    L.sunny_geneig_varlen_set_graphs.argtypes = [vp, ci]
    L.sunny_geneig_varlen_set_graphs.restype = ci
    L.sunny_geneig_varlen_graph_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong),
                                                  ctypes.POINTER(ctypes.c_longlong)]
    L.sunny_geneig_varlen_graph_stats.restype = ci
    L.sunny_geneig_varlen_num_buckets.argtypes = [vp]
    L.sunny_geneig_varlen_num_buckets.restype = ci
    L.sunny_geneig_varlen_bucket_order.argtypes = [vp, ci]
    L.sunny_geneig_varlen_bucket_order.restype = ci
    L.sunny_geneig_set_stage_timing.argtypes = [vp, ci]
    L.sunny_geneig_set_stage_timing.restype = ci
    L.sunny_geneig_set_graphs.argtypes = [vp, ci]
    L.sunny_geneig_set_graphs.restype = ci
    L.sunny_geneig_graph_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]
    L.sunny_geneig_graph_stats.restype = ci
    L.sunny_geneig_stage_times.argtypes = [vp, ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_char_p), ci]
    L.sunny_geneig_stage_times.restype = ci
    i64, u64 = ctypes.c_int64, ctypes.c_uint64
    pd = ctypes.POINTER(ctypes.c_double)
    L.sunny_geneig_session_create.argtypes = [ci, ci, pi, ci, ci, ci, u64, ci]
    L.sunny_geneig_session_create.restype = vp
    L.sunny_geneig_session_destroy.argtypes = [vp]
    L.sunny_geneig_session_destroy.restype = None
    L.sunny_geneig_session_submit.argtypes = [vp, ci, ci, cu, cu, cu, cu, cu, ci, cu]
    L.sunny_geneig_session_submit.restype = i64
    L.sunny_geneig_session_set_policy.argtypes = [vp, ci, ci]
    L.sunny_geneig_session_set_policy.restype = ci
    L.sunny_geneig_session_poll.argtypes = [vp, i64]
    L.sunny_geneig_session_poll.restype = ci
    L.sunny_geneig_session_wait.argtypes = [vp, i64, pd, pd]
    L.sunny_geneig_session_wait.restype = ci
    L.sunny_geneig_session_stats.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64),
                                             ctypes.POINTER(u64), pi]
    L.sunny_geneig_session_stats.restype = ci
    L.sunny_geneig_session_request_bytes.argtypes = [ci, ci]
    L.sunny_geneig_session_request_bytes.restype = u64
    _lib = L
    return L


def last_error() -> str:
    return lib().sunny_geneig_last_error().decode()


class GeneigError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"status {status}: {msg}")
        self.status = status


class Solver:
    """Solver(n; device_id=0, bpb_request=0, tuning='suggested', cache='auto',
    kernel='unified') -- SunnyGeneigBench.jl:37-57. kernel is 'unified' (fused
    FP64), 'hybrid' (split FP64) or 'mixed' (FP32 stages + FP64 refinement)."""

    def __init__(self, n: int, device_id: int = 0, bpb_request: int = 0,
                 tuning: str = "suggested", cache: str = "auto", kernel: str = "unified"):
        L = lib()
        t = TUNING_AUTOTUNE if tuning == "autotune" else TUNING_SUGGESTED
        c = {"nodisk": CACHE_NODISK, "nocache": CACHE_NOCACHE}.get(cache, CACHE_AUTO)
        k = {"hybrid": KERNEL_HYBRID, "mixed": KERNEL_MIXED}.get(kernel, KERNEL_UNIFIED)
        h = L.sunny_geneig_create_ex(n, device_id, bpb_request, t, c, k)
        if not h:
            raise GeneigError(STATUS_INVALID_ARG, f"sunny_geneig_create_ex failed: {last_error()}")
        self.handle = h
        self.n = n
        self.kernel = kernel
        self.device_id = device_id

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().sunny_geneig_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # Introspection (SunnyGeneigBench.jl:69-92).
    @property
    def block_dim_x(self) -> int:
        return lib().sunny_geneig_block_dim_x(self.handle)

    @property
    def batches_per_block(self) -> int:
        return lib().sunny_geneig_batches_per_block(self.handle)

    @property
    def shared_mem_bytes(self) -> int:
        return lib().sunny_geneig_shared_mem_bytes(self.handle)

    @property
    def path_name(self) -> str:
        return lib().sunny_geneig_path_name(self.handle).decode()

    @property
    def kernel_mode(self) -> str:
        return {KERNEL_HYBRID: "hybrid", KERNEL_MIXED: "mixed"}.get(
            lib().sunny_geneig_kernel_mode(self.handle), "unified")

    @property
    def staged(self) -> bool:
        """True when the handle runs the HBM-staged pipeline (orders above 64,
        or the autotuner's choice below)."""
        return lib().sunny_geneig_blocking(self.handle) == 1

    @property
    def cache_layer(self) -> str:
        return lib().sunny_geneig_cache_layer(self.handle).decode()

    def launch_raw(self, d_H: int, d_S: int, d_W: int, d_U: int, d_info: int, batch: int,
                   stream: int = 0) -> None:
        st = lib().sunny_geneig_launch(self.handle, d_H, d_S, d_W, d_U, d_info, batch, stream)
        if st != STATUS_OK:
            raise GeneigError(st, last_error())

    def set_graphs(self, on: bool = True) -> None:
        """Replay captured hipGraphs for repeated multi-kernel launches (default on)."""
        st = lib().sunny_geneig_set_graphs(self.handle, 1 if on else 0)
        if st != STATUS_OK:
            raise GeneigError(st, last_error())

    def graph_stats(self):
        """(captures, replays) of this handle's hipGraph cache."""
        c, r = ctypes.c_longlong(), ctypes.c_longlong()
        st = lib().sunny_geneig_graph_stats(self.handle, ctypes.byref(c), ctypes.byref(r))
        if st != STATUS_OK:
            raise GeneigError(st, last_error())
        return c.value, r.value

    def set_stage_timing(self, on: bool = True) -> None:
        st = lib().sunny_geneig_set_stage_timing(self.handle, 1 if on else 0)
        if st != STATUS_OK:
            raise GeneigError(st, last_error())

    def stage_times(self):
        """[(kernel name, ms)] of the last launch (stage timing must be on)."""
        ms = (ctypes.c_double * 12)()
        names = (ctypes.c_char_p * 12)()
        n = lib().sunny_geneig_stage_times(self.handle, ms, names, 12)
        if n < 0:
            raise GeneigError(STATUS_INVALID_ARG, last_error())
        return [(names[i].decode(), ms[i]) for i in range(min(n, 12))]


def _stream_handle(stream) -> int:
    if stream is None:
        import torch
        return int(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return stream
    return int(stream.cuda_stream)


def _check_tensor(name, t, shape, dtype, device):
    """The positional checks of run_kernel! (SunnyGeneigBench.jl:152-156):
    shape, element type, device and a dense layout the kernels can index."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
    if t.dtype != dtype:
        raise ValueError(f"{name} has dtype {t.dtype}, expected {dtype}")
    if t.device != device:
        raise ValueError(f"{name} is on {t.device}, expected {device}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


def _check_batch(solver: Solver, H, S, W, U, info):
    """Validate H, S (and W, U, info when given); returns (batch, device)."""
    import torch
    if H.dim() != 3:
        raise ValueError(f"H must be (batch, n, n), got shape {tuple(H.shape)}")
    batch, n, n2 = H.shape
    if n != solver.n or n2 != solver.n:
        raise ValueError(f"matrix size mismatch: solver.n={solver.n}, H={tuple(H.shape)}")
    dev = H.device
    if not H.is_cuda:
        raise ValueError("H must be a device tensor")
    _check_tensor("H", H, (batch, n, n), torch.complex128, dev)
    _check_tensor("S", S, (batch, n, n), torch.complex128, dev)
    if W is not None:
        _check_tensor("W", W, (batch, n), torch.float64, dev)
    if U is not None:
        _check_tensor("U", U, (batch, n, n), torch.complex128, dev)
    if info is not None:
        _check_tensor("info", info, (batch,), torch.int32, dev)
    return batch, dev


def run_kernel(solver: Solver, H, S, W=None, U=None, info=None, stream=None):
    """run_kernel!(solver, H, S; W, U, info, stream) -- SunnyGeneigBench.jl:104-171.

    H, S: complex128 cuda tensors (B, n, n) in column-major-per-matrix layout.
    Returns (W, U, info); pass output tensors to reuse them across calls
    (W float64 (B, n), U complex128 (B, n, n), info int32 (B,), contiguous, on
    H's device; anything else raises ValueError before the launch)."""
    import torch
    batch, dev = _check_batch(solver, H, S, W, U, info)
    n = solver.n
    W = torch.empty((batch, n), dtype=torch.float64, device=dev) if W is None else W
    U = torch.empty((batch, n, n), dtype=torch.complex128, device=dev) if U is None else U
    info = torch.zeros((batch,), dtype=torch.int32, device=dev) if info is None else info
    solver.launch_raw(H.data_ptr(), S.data_ptr(), W.data_ptr(), U.data_ptr(), info.data_ptr(),
                      batch, _stream_handle(stream))
    return W, U, info


def run_kernel_hdiag(solver: Solver, hdiag, S, W=None, U=None, info=None, stream=None):
    """run_kernel for a real diagonal H (Sunny's H = Sigma = diag(I, -I)):
    hdiag is float64 on S's device, shape (n,) (one diagonal for the batch) or
    (B, n) (one per matrix). Returns (W, U, info) as run_kernel."""
    import torch
    batch, dev = _check_batch(solver, S, S, W, U, info)
    n = solver.n
    if hdiag.dtype != torch.float64 or hdiag.device != dev or not hdiag.is_contiguous():
        raise ValueError("hdiag must be a contiguous float64 tensor on S's device")
    if tuple(hdiag.shape) == (n,):
        stride = 0
    elif tuple(hdiag.shape) == (batch, n):
        stride = n
    else:
        raise ValueError(f"hdiag must have shape ({n},) or ({batch}, {n}), got {tuple(hdiag.shape)}")
    W = torch.empty((batch, n), dtype=torch.float64, device=dev) if W is None else W
    U = torch.empty((batch, n, n), dtype=torch.complex128, device=dev) if U is None else U
    info = torch.zeros((batch,), dtype=torch.int32, device=dev) if info is None else info
    st = lib().sunny_geneig_launch_hdiag(solver.handle, hdiag.data_ptr(), stride, S.data_ptr(), W.data_ptr(),
                                         U.data_ptr(), info.data_ptr(), batch, _stream_handle(stream))
    if st != STATUS_OK:
        raise GeneigError(st, last_error())
    return W, U, info


def run_kernel_chunked(solver: Solver, H, S, W, U, info, chunk_size: int, streams):
    """run_kernel_chunked! (SunnyGeneigBench.jl:191-251): contiguous chunks of
    chunk_size matrices round-robin over `streams` through
    sunny_geneig_launch_chunked, then every stream is synchronised."""
    batch, _ = _check_batch(solver, H, S, W, U, info)
    if W is None or U is None or info is None:
        raise ValueError("run_kernel_chunked needs preallocated W, U and info")
    if chunk_size <= 0 or not streams:
        raise ValueError("chunk_size must be > 0 and at least one stream given")
    handles = (ctypes.c_uint64 * len(streams))(*[_stream_handle(st) for st in streams])
    st = lib().sunny_geneig_launch_chunked(solver.handle, H.data_ptr(), S.data_ptr(), W.data_ptr(),
                                           U.data_ptr(), info.data_ptr(), batch, chunk_size,
                                           handles, len(streams))
    if st != STATUS_OK:
        raise GeneigError(st, last_error())
    for s_ in streams:
        s_.synchronize()
    return W, U, info


def precompile(n: int, block_dim_x: int = 0, arch: str = "gfx950") -> None:
    st = lib().sunny_geneig_precompile(n, block_dim_x, arch.encode())
    if st != STATUS_OK:
        raise GeneigError(st, last_error())


def cache_dir() -> str:
    return lib().sunny_geneig_cache_dir().decode()


class VarlenSolver:
    """Uneven batches: one call solves pencils of different orders n_k <= max_n.

    Matrices are packed back to back (matrix k column-major at element offset
    sum_{j<k} n_j^2; eigenvalues at sum_{j<k} n_j); `pack` / `unpack` convert
    from and to lists of (n_k, n_k) tensors in the (B, n, n) convention above."""

    def __init__(self, max_n: int = 64, device_id: int = 0, buckets=None, cache: str = "auto",
                 tuning: str = "suggested"):
        """tuning: "suggested" (fixed workgroup size per order) or "autotune"
        (per bucket, the split-mode solver's timed sweep for that order)."""
        L = lib()
        c = {"nodisk": CACHE_NODISK, "nocache": CACHE_NOCACHE}.get(cache, CACHE_AUTO)
        if tuning not in ("suggested", "autotune"):
            raise ValueError("tuning must be 'suggested' or 'autotune'")
        t = TUNING_AUTOTUNE if tuning == "autotune" else TUNING_SUGGESTED
        if buckets:
            arr = (ctypes.c_int * len(buckets))(*buckets)
            h = L.sunny_geneig_varlen_create_ex(max_n, device_id, arr, len(buckets), c, t)
        else:
            h = L.sunny_geneig_varlen_create_ex(max_n, device_id, None, 0, c, t)
        if not h:
            raise GeneigError(STATUS_INVALID_ARG, f"sunny_geneig_varlen_create failed: {last_error()}")
        self.handle = h
        self.max_n = max_n

    @property
    def buckets(self):
        L = lib()
        return [L.sunny_geneig_varlen_bucket_order(self.handle, i)
                for i in range(L.sunny_geneig_varlen_num_buckets(self.handle))]

    def set_graphs(self, on: bool = True) -> None:
        """Capture each call into a hipGraph and replay it when the buffers
        and orders repeat (default on)."""
        st = lib().sunny_geneig_varlen_set_graphs(self.handle, 1 if on else 0)
        if st != STATUS_OK:
            raise GeneigError(st, last_error())

    def graph_stats(self):
        """(captures, replays) of this handle's call graphs."""
        c, r = ctypes.c_longlong(), ctypes.c_longlong()
        st = lib().sunny_geneig_varlen_graph_stats(self.handle, ctypes.byref(c), ctypes.byref(r))
        if st != STATUS_OK:
            raise GeneigError(st, last_error())
        return c.value, r.value

    @property
    def bucket_waves(self):
        """Matrices per workgroup of each bucket's grid."""
        L = lib()
        return [L.sunny_geneig_varlen_bucket_waves(self.handle, i)
                for i in range(L.sunny_geneig_varlen_num_buckets(self.handle))]

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().sunny_geneig_varlen_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def launch_raw(self, d_H: int, d_S: int, d_W: int, d_U: int, d_info: int, orders,
                   stream: int = 0) -> None:
        arr = (ctypes.c_int * len(orders))(*[int(x) for x in orders])
        st = lib().sunny_geneig_varlen_launch(self.handle, d_H, d_S, d_W, d_U, d_info, arr,
                                              len(orders), stream)
        if st != STATUS_OK:
            raise GeneigError(st, last_error())

    def run(self, H, S, orders, W=None, U=None, info=None, stream=None):
        """H, S: packed complex128 device tensors (sum n_k^2,). Returns packed
        (W, U, info)."""
        import torch
        orders = [int(x) for x in orders]
        nsq = sum(x * x for x in orders)
        if H.numel() != nsq or S.numel() != nsq:
            raise ValueError(f"packed H/S must hold sum n_k^2 = {nsq} elements")
        for name, t in (("H", H), ("S", S)):
            if t.dtype != torch.complex128 or not t.is_cuda or not t.is_contiguous():
                raise ValueError(f"{name} must be a contiguous complex128 device tensor")
        dev = H.device
        W = torch.empty((sum(orders),), dtype=torch.float64, device=dev) if W is None else W
        U = torch.empty((nsq,), dtype=torch.complex128, device=dev) if U is None else U
        info = torch.zeros((len(orders),), dtype=torch.int32, device=dev) if info is None else info
        self.launch_raw(H.data_ptr(), S.data_ptr(), W.data_ptr(), U.data_ptr(), info.data_ptr(),
                        orders, _stream_handle(stream))
        return W, U, info


def pack(mats):
    """List of (n_k, n_k) tensors -> one flat packed tensor."""
    import torch
    return torch.cat([m.reshape(-1) for m in mats]).contiguous()


def unpack_matrices(flat, orders):
    out, o = [], 0
    for n in orders:
        out.append(flat[o:o + n * n].view(n, n))
        o += n * n
    return out


def unpack_vectors(flat, orders):
    out, o = [], 0
    for n in orders:
        out.append(flat[o:o + n])
        o += n
    return out


class ServingSession:
    """Request loop over resident solvers (sunny_geneig_session_*).

    ServingSession(orders=[24, 54], kernel="mixed", max_count=4096, slots=4,
    hbm_budget=2 << 30, policy="fifo", slice=0). submit(H, S, W, U, info)
    queues one request: with device tensors the buffers are used in place (the
    session waits for the current torch stream as it is at submission), with
    (pinned) CPU tensors the session stages them through its device arena.
    policy "fifo" serves requests in arrival order, "srpt" the one with the
    least work left; slice > 0 cuts requests into slices of that many pencils.
    Returns a request id; wait(id) -> (latency_ms, service_ms)."""

    POLICIES = {"fifo": 0, "srpt": 1}

    def __init__(self, orders, kernel: str = "mixed", max_count: int = 4096, slots: int = 4,
                 hbm_budget: int = 2 << 30, device_id: int = 0, cache: str = "auto",
                 policy: str = "fifo", slice: int = 0):
        L = lib()
        k = {"hybrid": KERNEL_HYBRID, "mixed": KERNEL_MIXED}.get(kernel, KERNEL_UNIFIED)
        c = {"nodisk": CACHE_NODISK, "nocache": CACHE_NOCACHE}.get(cache, CACHE_AUTO)
        arr = (ctypes.c_int * len(orders))(*orders)
        h = L.sunny_geneig_session_create(device_id, k, arr, len(orders), max_count, slots,
                                          hbm_budget, c)
        if not h:
            raise GeneigError(STATUS_INVALID_ARG, f"sunny_geneig_session_create failed: {last_error()}")
        self.handle = h
        self.device_id = device_id
        self._keep = {}
        self.set_policy(policy, slice)

    def set_policy(self, policy: str = "fifo", slice: int = 0) -> None:
        if policy not in self.POLICIES:
            raise ValueError(f"policy must be one of {sorted(self.POLICIES)}")
        st = lib().sunny_geneig_session_set_policy(self.handle, self.POLICIES[policy], int(slice))
        if st != STATUS_OK:
            raise GeneigError(st, last_error())

    def submit(self, H, S, W, U, info) -> int:
        """Queue one request of K pencils: H, S complex128 (K, n, n), W float64
        (K, n), U complex128 (K, n, n), info int32 (K,), all contiguous and all
        on the session's device or all in host memory."""
        import torch
        if H.dim() != 3 or H.shape[1] != H.shape[2]:
            raise ValueError(f"H must be (count, n, n), got shape {tuple(H.shape)}")
        count, n = H.shape[0], H.shape[-1]
        on_dev = H.is_cuda
        dev = H.device
        if on_dev and dev.index not in (None, self.device_id):
            raise ValueError(f"H is on {dev}, the session serves cuda:{self.device_id}")
        for name, t, shape, dtype in (("H", H, (count, n, n), torch.complex128),
                                      ("S", S, (count, n, n), torch.complex128),
                                      ("W", W, (count, n), torch.float64),
                                      ("U", U, (count, n, n), torch.complex128),
                                      ("info", info, (count,), torch.int32)):
            _check_tensor(name, t, shape, dtype, dev)
        stream = int(torch.cuda.current_stream(dev).cuda_stream) if on_dev else 0
        rid = lib().sunny_geneig_session_submit(self.handle, n, count, H.data_ptr(), S.data_ptr(),
                                                W.data_ptr(), U.data_ptr(), info.data_ptr(),
                                                1 if on_dev else 0, stream)
        if rid < 0:
            raise GeneigError(int(rid), last_error())
        self._keep[rid] = (H, S, W, U, info)     # buffers stay alive until wait()
        return rid

    def poll(self, rid: int) -> bool:
        st = lib().sunny_geneig_session_poll(self.handle, rid)
        if st < 0:
            raise GeneigError(st, last_error())
        return st == 1

    def wait(self, rid: int):
        lat, svc = ctypes.c_double(), ctypes.c_double()
        st = lib().sunny_geneig_session_wait(self.handle, rid, ctypes.byref(lat), ctypes.byref(svc))
        self._keep.pop(rid, None)
        if st != STATUS_OK:
            raise GeneigError(st, last_error())
        return lat.value, svc.value

    def stats(self):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        c, d = ctypes.c_uint64(), ctypes.c_int()
        lib().sunny_geneig_session_stats(self.handle, ctypes.byref(a), ctypes.byref(b),
                                         ctypes.byref(c), ctypes.byref(d))
        return {"submitted": a.value, "completed": b.value, "hbm_peak_bytes": c.value,
                "max_concurrent": d.value}

    @staticmethod
    def request_bytes(n: int, count: int) -> int:
        return lib().sunny_geneig_session_request_bytes(n, count)

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().sunny_geneig_session_destroy(self.handle)
            self.handle = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
