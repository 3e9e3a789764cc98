"""Synthetic input batches generated on the device with torch.

The benchmark's input family (H = Q diag(lambda) Q^H, Q Haar-unitary,
lambda ~ U[-10, 10] with `n_pairs` near-degenerate pairs split by `pair_gap`;
S = B B^H + 0.05 N I with B complex Gaussian, re and im ~ N(0, 1) each) or,
with cond_s > 0, S = Q_S diag(sigma) Q_S^H with sigma geometric in [1, cond_s].

Q comes from the product of N random Householder reflectors (G. W. Stewart,
"The efficient generation of random orthogonal matrices with an application
to condition estimators", SIAM J. Numer. Anal. 17 (1980)), which costs N
batched rank-1 updates on the GPU instead of a batched QR.

Outputs use the library's layout: (B, n, n) complex128 whose [b, j, i] is
element (i, j) (column-major per matrix).
"""
from __future__ import annotations

import torch


def _haar_unitary(batch: int, n: int, gen: torch.Generator, device) -> torch.Tensor:
    Q = torch.eye(n, dtype=torch.complex128, device=device).repeat(batch, 1, 1)
    for k in range(n):
        m = n - k
        x = torch.randn((batch, m), dtype=torch.complex128, generator=gen, device=device)
        nx = torch.linalg.vector_norm(x, dim=1)
        a0 = x[:, 0]
        ph = torch.where(a0.abs() > 0, a0 / a0.abs().clamp_min(1e-300), torch.ones_like(a0))
        v = x.clone()
        v[:, 0] = v[:, 0] + ph * nx
        v = v / torch.linalg.vector_norm(v, dim=1, keepdim=True)
        Qk = Q[:, :, k:].contiguous()
        w = torch.bmm(Qk, v.unsqueeze(-1))                        # (B, n, 1)
        Q[:, :, k:] = Qk - 2.0 * torch.bmm(w, v.conj().unsqueeze(1).contiguous())
    return Q


def _colmajor(A: torch.Tensor) -> torch.Tensor:
    return A.transpose(-1, -2).contiguous()


def _hermitian_from(Q: torch.Tensor, w: torch.Tensor) -> torch.Tensor:
    X = torch.bmm(Q * w.unsqueeze(1).to(Q.dtype), Q.conj().transpose(-1, -2).contiguous())
    return 0.5 * (X + X.conj().transpose(-1, -2))


@torch.no_grad()
def spectrum_batch(n: int, batch: int, seed: int, device="cuda", n_pairs: int = 6,
                   pair_gap: float = 1e-8, cond_s=0.0, chunk: int = 8192):
    """Generate (H, S) on `device`; deterministic for a given seed and device.
    cond_s > 0 (a number, or one value per matrix) replaces S by
    Q_S diag(sigma) Q_S^H with sigma geometric from 1 to cond_s."""
    dev = torch.device(device)
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed) & 0x7FFFFFFFFFFFFFFF)
    H = torch.empty((batch, n, n), dtype=torch.complex128, device=dev)
    S = torch.empty((batch, n, n), dtype=torch.complex128, device=dev)
    eye = torch.eye(n, dtype=torch.complex128, device=dev)
    for s in range(0, batch, chunk):
        b = min(chunk, batch - s)
        Q = _haar_unitary(b, n, gen, dev)
        lam = torch.rand((b, n), dtype=torch.float64, generator=gen, device=dev) * 20.0 - 10.0
        for k in range(min(n_pairs, n // 2)):
            lam[:, 2 * k + 1] = lam[:, 2 * k] + pair_gap
        H[s:s + b] = _colmajor(_hermitian_from(Q, lam))
        if torch.is_tensor(cond_s) or cond_s > 0:
            QS = _haar_unitary(b, n, gen, dev)
            cs = torch.as_tensor(cond_s, dtype=torch.float64, device=dev).expand(batch)[s:s + b]
            frac = torch.arange(n, dtype=torch.float64, device=dev) / max(n - 1, 1)
            sig = torch.exp(torch.log(cs)[:, None] * frac[None, :])
            S[s:s + b] = _colmajor(_hermitian_from(QS, sig))
        else:
            Bm = torch.randn((b, n, n), dtype=torch.complex128, generator=gen, device=dev)
            # torch's complex randn has E|z|^2 = 1; scale to re, im ~ N(0, 1) each.
            Bm = Bm * (2.0 ** 0.5)
            X = torch.bmm(Bm, Bm.conj().transpose(-1, -2).contiguous()) + 0.05 * n * eye
            S[s:s + b] = _colmajor(0.5 * (X + X.conj().transpose(-1, -2)))
    return H, S


@torch.no_grad()
def quality(H: torch.Tensor, S: torch.Tensor, W: torch.Tensor, U: torch.Tensor, chunk: int = 8192):
    """Per-matrix residual ||H U - S U diag(W)||_1 / (||H||_1 ||U||_1) and
    S-orthogonality ||U^H S U - I||_1, computed on the device (float64)."""
    n = H.shape[-1]
    res, orth = [], []
    eye = torch.eye(n, dtype=torch.complex128, device=H.device)
    for s in range(0, H.shape[0], chunk):
        # Column-major storage: the row-major view of each matrix is its transpose.
        h = H[s:s + chunk].transpose(-1, -2)
        sm = S[s:s + chunk].transpose(-1, -2)
        u = U[s:s + chunk].transpose(-1, -2).contiguous()
        w = W[s:s + chunk]
        hu = torch.bmm(h.contiguous(), u)
        su = torch.bmm(sm.contiguous(), u)
        r = hu - su * w.unsqueeze(1).to(su.dtype)
        n1 = lambda x: x.abs().sum(dim=-2).amax(dim=-1)  # noqa: E731  (max column sum)
        res.append(n1(r) / (n1(h) * n1(u)))
        g = torch.bmm(u.conj().transpose(-1, -2).contiguous(), su) - eye
        orth.append(n1(g))
    return torch.cat(res), torch.cat(orth)
