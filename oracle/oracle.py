"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker. The product path never does.

Arrays are numpy, column-major per matrix: a batch of B matrices of order N is
an ndarray of shape (B, N, N) whose [b] slice is the Fortran-ordered matrix,
i.e. the flat memory is b*N*N + j*N + i for element (i, j).
"""
from __future__ import annotations

import ctypes
import glob
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libgeneig_oracle.so")
_lib = None

_c128p = ctypes.POINTER(ctypes.c_double)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)


def build() -> None:
    """Compile the oracle with its Makefile (g++)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(_i32p)


def lapack_candidates():
    """LAPACK providers to try, most specific first."""
    cands = []
    env = os.environ.get("GENEIG_LAPACK_LIB")
    if env:
        cands.append(env)
    try:
        import scipy  # noqa: F401  (only to locate its bundled OpenBLAS)
        d = os.path.join(os.path.dirname(scipy.__file__), "..", "scipy.libs")
        cands += sorted(glob.glob(os.path.join(d, "libscipy_openblas*.so")))
    except Exception:
        pass
    cands += ["liblapack.so.3", "libopenblas.so.0"]
    return cands


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = ctypes.CDLL(_LIB_PATH)
    L.go_gen_reference.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_int,
                                   _c128p, _c128p]
    L.go_gen_spectrum.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_double, _c128p, _c128p]
    L.go_geig_read_header.argtypes = [ctypes.c_char_p, _i32p]
    L.go_geig_read.argtypes = [ctypes.c_char_p, _c128p, _c128p, _f64p, _c128p]
    L.go_geig_write.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                _c128p, _c128p, _f64p, _c128p]
    L.go_lapack_load.argtypes = [ctypes.c_char_p]
    L.go_lapack_provider.restype = ctypes.c_char_p
    L.go_zhegv.argtypes = [ctypes.c_int, _c128p, _c128p, _f64p, _c128p]
    L.go_zhegv_batch.argtypes = [ctypes.c_int, ctypes.c_int, _c128p, _c128p, _f64p, _c128p,
                                 _i32p, ctypes.c_int]
    L.go_zpotrf.argtypes = [ctypes.c_int, _c128p]
    L.go_zhegst.argtypes = [ctypes.c_int, _c128p, _c128p]
    L.go_zheev.argtypes = [ctypes.c_int, _c128p, _f64p]
    L.go_dsteqr.argtypes = [ctypes.c_int, _f64p, _f64p, _f64p]
    L.go_rs_potrf.argtypes = [ctypes.c_int, _c128p]
    L.go_rs_reduce.argtypes = [ctypes.c_int, _c128p, _c128p]
    L.go_rs_hetrd.argtypes = [ctypes.c_int, _c128p, _f64p, _f64p, _c128p]
    L.go_rs_tql.argtypes = [ctypes.c_int, _f64p, _f64p, _f64p]
    L.go_rs_apply_q.argtypes = [ctypes.c_int, _c128p, _c128p, _c128p]
    L.go_rs_backtrsm.argtypes = [ctypes.c_int, _c128p, _c128p]
    L.go_rs_zhegv.argtypes = [ctypes.c_int, _c128p, _c128p, _f64p, _c128p]
    L.go_rs_zhegv_batch.argtypes = [ctypes.c_int, ctypes.c_int, _c128p, _c128p, _f64p, _c128p,
                                    _i32p, ctypes.c_int]
    L.go_compare.argtypes = [ctypes.c_int, _c128p, _f64p, _c128p, _f64p, _c128p,
                             ctypes.c_double, _f64p]
    L.go_residual.argtypes = [ctypes.c_int, _c128p, _c128p, _f64p, _c128p]
    L.go_residual.restype = ctypes.c_double
    L.go_s_orth.argtypes = [ctypes.c_int, _c128p, _c128p]
    L.go_s_orth.restype = ctypes.c_double
    L.go_check_batch.argtypes = [ctypes.c_int, ctypes.c_int, _c128p, _c128p, _f64p, _c128p,
                                 _f64p, _c128p, ctypes.c_double, ctypes.c_int, _f64p]
    _lib = L
    return L


_lapack_loaded = False


def load_lapack() -> str:
    """Load a LAPACK provider; returns its description. Raises if none."""
    global _lapack_loaded
    L = lib()
    if _lapack_loaded:
        return L.go_lapack_provider().decode()
    for c in lapack_candidates():
        if L.go_lapack_load(c.encode()) == 0:
            _lapack_loaded = True
            return L.go_lapack_provider().decode()
    raise RuntimeError("no LAPACK provider found (tried %s)" % lapack_candidates())


def _cbatch(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a


def _empty_c(batch, n):
    return np.zeros((batch, n, n), dtype=np.complex128)


def to_fortran_batch(a: np.ndarray) -> np.ndarray:
    """(B, N, N) numpy matrices (row-major semantics) -> flat column-major layout."""
    return np.ascontiguousarray(np.swapaxes(a, -1, -2))


def from_fortran_batch(a: np.ndarray) -> np.ndarray:
    """Inverse of to_fortran_batch: column-major storage -> (B, N, N) matrices."""
    return np.ascontiguousarray(np.swapaxes(a, -1, -2))


# ---- generators --------------------------------------------------------------

def gen_reference(n: int, batch: int, seed: int, s_mode: str = "random"):
    """Reference fixture generator (column-major storage, shape (B, N, N))."""
    H = _empty_c(batch, n)
    S = _empty_c(batch, n)
    lib().go_gen_reference(n, batch, seed, 0 if s_mode == "random" else 1, _ptr(H), _ptr(S))
    return H, S


def gen_spectrum(n: int, batch: int, seed: int, n_pairs: int = 6, pair_gap: float = 1e-8,
                 cond_s: float = 0.0):
    H = _empty_c(batch, n)
    S = _empty_c(batch, n)
    lib().go_gen_spectrum(n, batch, seed, n_pairs, pair_gap, cond_s, _ptr(H), _ptr(S))
    return H, S


# ---- GEIG v1 -------------------------------------------------------------------

def geig_read(path: str):
    hdr = np.zeros(6, dtype=np.int32)
    rc = lib().go_geig_read_header(path.encode(), _iptr(hdr))
    if rc:
        raise ValueError(f"bad GEIG file {path}: rc={rc}")
    n, b = int(hdr[2]), int(hdr[3])
    H, S, U = _empty_c(b, n), _empty_c(b, n), _empty_c(b, n)
    W = np.zeros((b, n), dtype=np.float64)
    rc = lib().go_geig_read(path.encode(), _ptr(H), _ptr(S), _ptr(W), _ptr(U))
    if rc:
        raise ValueError(f"short GEIG file {path}: rc={rc}")
    return dict(n=n, batch=b, seed=int(hdr[4]), H=H, S=S, W=W, U=U)


def geig_write(path: str, n, batch, seed, H, S, W, U) -> None:
    rc = lib().go_geig_write(path.encode(), n, batch, seed, _ptr(_cbatch(H)), _ptr(_cbatch(S)),
                             _ptr(np.ascontiguousarray(W, dtype=np.float64)), _ptr(_cbatch(U)))
    if rc:
        raise IOError(f"GEIG write failed: {rc}")


# ---- LAPACK ---------------------------------------------------------------------

def zhegv_batch(H, S, threads: int = 1):
    load_lapack()
    H = _cbatch(H)
    S = _cbatch(S)
    b, n = H.shape[0], H.shape[1]
    W = np.zeros((b, n))
    U = _empty_c(b, n)
    info = np.zeros(b, dtype=np.int32)
    rc = lib().go_zhegv_batch(n, b, _ptr(H), _ptr(S), _ptr(W), _ptr(U), _iptr(info), threads)
    if rc:
        raise RuntimeError("LAPACK not loaded")
    return W, U, info


def zpotrf(A):
    load_lapack()
    A = _cbatch(A).copy()
    info = lib().go_zpotrf(A.shape[-1], _ptr(A))
    return A, info


def zhegst(A, L):
    load_lapack()
    A = _cbatch(A).copy()
    info = lib().go_zhegst(A.shape[-1], _ptr(A), _ptr(_cbatch(L)))
    return A, info


def zheev(A):
    load_lapack()
    A = _cbatch(A).copy()
    n = A.shape[-1]
    W = np.zeros(n)
    info = lib().go_zheev(n, _ptr(A), _ptr(W))
    return W, A, info


def dsteqr(d, e):
    load_lapack()
    d = np.array(d, dtype=np.float64)
    n = d.shape[0]
    e2 = np.zeros(max(n, 1))
    e2[: n - 1] = e[: n - 1]
    Z = np.zeros((n, n))
    info = lib().go_dsteqr(n, _ptr(d), _ptr(e2), _ptr(Z))
    return d, Z, info


# ---- restatement ------------------------------------------------------------------

def rs_potrf(A):
    A = _cbatch(A).copy()
    info = lib().go_rs_potrf(A.shape[-1], _ptr(A))
    return A, info


def rs_reduce(L, A):
    A = _cbatch(A).copy()
    lib().go_rs_reduce(A.shape[-1], _ptr(_cbatch(L)), _ptr(A))
    return A


def rs_hetrd(A):
    A = _cbatch(A).copy()
    n = A.shape[-1]
    d, e = np.zeros(n), np.zeros(n)
    tau = np.zeros(n, dtype=np.complex128)
    lib().go_rs_hetrd(n, _ptr(A), _ptr(d), _ptr(e), _ptr(tau))
    return A, d, e, tau


def rs_tql(d, e):
    d = np.array(d, dtype=np.float64)
    n = d.shape[0]
    e2 = np.zeros(n)
    e2[: n - 1] = np.asarray(e)[: n - 1]
    Z = np.zeros((n, n))
    info = lib().go_rs_tql(n, _ptr(d), _ptr(e2), _ptr(Z))
    return d, Z, info


def rs_apply_q(A, tau, V):
    V = _cbatch(V).copy()
    lib().go_rs_apply_q(V.shape[-1], _ptr(_cbatch(A)), _ptr(_cbatch(tau)), _ptr(V))
    return V


def rs_backtrsm(L, V):
    V = _cbatch(V).copy()
    lib().go_rs_backtrsm(V.shape[-1], _ptr(_cbatch(L)), _ptr(V))
    return V


def rs_zhegv_batch(H, S, threads: int = 1):
    H = _cbatch(H)
    S = _cbatch(S)
    b, n = H.shape[0], H.shape[1]
    W = np.zeros((b, n))
    U = _empty_c(b, n)
    info = np.zeros(b, dtype=np.int32)
    lib().go_rs_zhegv_batch(n, b, _ptr(H), _ptr(S), _ptr(W), _ptr(U), _iptr(info), threads)
    return W, U, info


# ---- metrics ------------------------------------------------------------------------

def check_batch(H, S, W, U, W_ref=None, U_ref=None, degen_gap: float = 1e-6,
                threads: int = 8):
    """Per-matrix metrics, shape (B, 5): eig_rel, phase_err, n_excluded,
    residual (1-norm), S-orthogonality (1-norm)."""
    H, S, U = _cbatch(H), _cbatch(S), _cbatch(U)
    W = np.ascontiguousarray(W, dtype=np.float64)
    b, n = H.shape[0], H.shape[1]
    out = np.zeros((b, 5))
    wr = _ptr(np.ascontiguousarray(W_ref, dtype=np.float64)) if W_ref is not None else None
    ur = _ptr(_cbatch(U_ref)) if U_ref is not None else None
    lib().go_check_batch(n, b, _ptr(H), _ptr(S), wr, ur, _ptr(W), _ptr(U), degen_gap, threads,
                         _ptr(out))
    return out
