"""GEIG on-disk format for batches of generalized eigenproblems.

Version 1 is the reference's fixture format (tests/common/lapack_reference.cpp:257-279,
read by tests/common/dump_pair.cpp:30-114): an int32[6] header
{magic 0x47454947 ("GEIG"), 1, N, B, seed, 0}, then H and S (complex128,
B x N x N), W (float64, B x N) and U (complex128, B x N x N), every matrix
column-major, little-endian.

Version 2 (no counterpart in the reference) holds uneven batches and
per-pencil status: an int32[6] header {magic, 2, count, seed, flags, 0},
int32 orders[count], then H and S packed back to back (pencil k column-major
at element offset sum_{j<k} n_j^2, complex128), and, when flags bit 0 is set,
W (float64, sum n_k, pencil k at sum_{j<k} n_j), U (packed like H) and
info (int32[count], zhegv's INFO). The packing is the uneven-batch path's
(`VarlenSolver`), so a file feeds it without reshuffling.

Arrays use the library's layout: a (B, N, N) array `A[b, j, i]` is element
(i, j) of matrix b (the memory of the reference's (N, N, B) arrays)."""
from __future__ import annotations

import numpy as np

MAGIC = 0x47454947
_HDR = np.dtype("<i4")


def write_v1(path, seed: int, H, S, W, U) -> None:
    """Write a version-1 file: H, S, U (B, N, N) complex128, W (B, N) float64."""
    H, S, U = (np.ascontiguousarray(a, dtype="<c16") for a in (H, S, U))
    W = np.ascontiguousarray(W, dtype="<f8")
    b, n = H.shape[0], H.shape[1]
    for name, a, shape in (("H", H, (b, n, n)), ("S", S, (b, n, n)), ("U", U, (b, n, n)), ("W", W, (b, n))):
        if a.shape != shape:
            raise ValueError(f"{name} has shape {a.shape}, expected {shape}")
    with open(path, "wb") as f:
        f.write(np.array([MAGIC, 1, n, b, seed, 0], dtype=_HDR).tobytes())
        for a in (H, S, W, U):
            f.write(a.tobytes())


def write_v2(path, seed: int, orders, H, S, W=None, U=None, info=None) -> None:
    """Write a version-2 file. H, S (and U): packed complex128 (sum n_k^2,);
    W: float64 (sum n_k,); info: int32 (count,). W, U and info go together."""
    orders = np.asarray(orders, dtype="<i4")
    if orders.ndim != 1 or (orders < 1).any():
        raise ValueError("orders must be a 1-D array of positive orders")
    nsq, nsum = int((orders.astype(np.int64) ** 2).sum()), int(orders.sum())
    H = np.ascontiguousarray(H, dtype="<c16").reshape(-1)
    S = np.ascontiguousarray(S, dtype="<c16").reshape(-1)
    if H.size != nsq or S.size != nsq:
        raise ValueError(f"packed H and S must hold sum n_k^2 = {nsq} elements")
    results = W is not None or U is not None or info is not None
    if results:
        if W is None or U is None or info is None:
            raise ValueError("W, U and info are written together")
        W = np.ascontiguousarray(W, dtype="<f8").reshape(-1)
        U = np.ascontiguousarray(U, dtype="<c16").reshape(-1)
        info = np.ascontiguousarray(info, dtype="<i4").reshape(-1)
        if W.size != nsum or U.size != nsq or info.size != orders.size:
            raise ValueError("W, U or info has the wrong length for these orders")
    with open(path, "wb") as f:
        f.write(np.array([MAGIC, 2, orders.size, seed, 1 if results else 0, 0], dtype=_HDR).tobytes())
        f.write(orders.tobytes())
        f.write(H.tobytes())
        f.write(S.tobytes())
        if results:
            for a in (W, U, info):
                f.write(a.tobytes())


def read(path) -> dict:
    """Read a version-1 or version-2 file. Version 1: {"version", "n", "batch",
    "seed", "H", "S", "W", "U"} with (B, N, N) / (B, N) arrays. Version 2:
    {"version", "seed", "orders", "H", "S"} plus "W", "U", "info" when stored,
    packed as written."""
    with open(path, "rb") as f:
        raw = f.read()
    if len(raw) < 24:
        raise ValueError(f"{path}: too short for a GEIG header")
    hdr = np.frombuffer(raw[:24], dtype=_HDR)
    if int(hdr[0]) != MAGIC:
        raise ValueError(f"{path}: bad magic {int(hdr[0]):#x}")
    version = int(hdr[1])
    off = 24

    def take(dtype, count):
        nonlocal off
        nbytes = np.dtype(dtype).itemsize * count
        if off + nbytes > len(raw):
            raise ValueError(f"{path}: truncated")
        a = np.frombuffer(raw[off:off + nbytes], dtype=dtype).copy()
        off += nbytes
        return a

    if version == 1:
        n, b, seed = int(hdr[2]), int(hdr[3]), int(hdr[4])
        m = n * n * b
        out = {"version": 1, "n": n, "batch": b, "seed": seed,
               "H": take("<c16", m).reshape(b, n, n), "S": take("<c16", m).reshape(b, n, n),
               "W": take("<f8", n * b).reshape(b, n), "U": take("<c16", m).reshape(b, n, n)}
    elif version == 2:
        count, seed, flags = int(hdr[2]), int(hdr[3]), int(hdr[4])
        orders = take("<i4", count)
        nsq, nsum = int((orders.astype(np.int64) ** 2).sum()), int(orders.sum())
        out = {"version": 2, "seed": seed, "orders": orders, "H": take("<c16", nsq), "S": take("<c16", nsq)}
        if flags & 1:
            out["W"] = take("<f8", nsum)
            out["U"] = take("<c16", nsq)
            out["info"] = take("<i4", count)
    else:
        raise ValueError(f"{path}: unsupported GEIG version {version}")
    if off != len(raw):
        raise ValueError(f"{path}: {len(raw) - off} trailing bytes")
    return out
