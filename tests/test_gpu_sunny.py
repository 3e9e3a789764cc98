"""The Sunny.jl call pattern (julia/SunnyGeneigBench/scripts/setup_sw08.jl:6-18):
H = Sigma = diag(I_L, -I_L), the paraunitary metric of Colpa's method, which is
indefinite, and S = D, the dynamical matrix, Hermitian positive definite with
the Bogoliubov block form D = [[A, B], [conj(B), conj(A)]] (A Hermitian, B
symmetric). Sigma D then has a spectrum of +-omega pairs, so the pencil
(Sigma, D) has eigenvalues in exact +- pairs: lambda_{2L-1-k} = -lambda_k.

Checked through the C ABI in every mode, at LDS-resident and staged orders:
the pairing (to 1e-12 relative), and parity with FP64 LAPACK zhegv on the
same pencils with the usual bars."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as T

pytestmark = pytest.mark.gpu


def bogoliubov_pencils(L: int, batch: int, seed: int):
    """(Sigma, D) pairs in the library layout (matrix b, element (i, j) at
    [b, j, i]); D = [[A, B], [B*, A*]] + shift I with A Hermitian, B symmetric,
    shifted to be positive definite with a condition number around 1e2."""
    rng = np.random.default_rng(seed)
    n = 2 * L
    H = np.zeros((batch, n, n), dtype=np.complex128)
    S = np.zeros((batch, n, n), dtype=np.complex128)
    sig = np.concatenate([np.ones(L), -np.ones(L)])
    for b in range(batch):
        X = rng.standard_normal((L, L)) + 1j * rng.standard_normal((L, L))
        A = X @ X.conj().T / L
        Y = rng.standard_normal((L, L)) + 1j * rng.standard_normal((L, L))
        Bs = 0.3 * (Y + Y.T) / 2
        D = np.block([[A, Bs], [Bs.conj(), A.conj()]])
        D = 0.5 * (D + D.conj().T)
        w = np.linalg.eigvalsh(D)
        D = D + (max(0.0, -w[0]) + 0.01 * (w[-1] - w[0]) + 1e-3) * np.eye(n)
        H[b] = np.diag(sig).T
        S[b] = D.T
    return H, S


@pytest.mark.parametrize("kernel", ["unified", "hybrid", "mixed"])
@pytest.mark.parametrize("L", [12, 27, 48])
def test_paraunitary_metric_pencil(kernel, L):
    H, S = bogoliubov_pencils(L, 8, 1000 + L)
    W, U, info = T.run_gpu(H, S, kernel=kernel)
    assert (info == 0).all()
    # +- pairing of the ascending eigenvalues.
    scale = np.abs(W).max(axis=1, keepdims=True)
    pair_err = np.abs(W + W[:, ::-1]) / scale
    assert pair_err.max() < 1e-12, f"pairing {pair_err.max():.2e}"
    Wl, Ul, il = O.zhegv_batch(H, S, threads=8)
    assert (il == 0).all()
    T.assert_parity(H, S, W, U, info, Wl, Ul, tag=f"{kernel} Sigma/D L={L}")


def _run_hdiag(H, S, kernel, per_matrix=False):
    """Solve through sunny_geneig_launch_hdiag with H's diagonal."""
    import torch
    import sunny_gen_eigensolver_amd as sg
    b, n = H.shape[0], H.shape[1]
    d = np.ascontiguousarray(np.real(np.diagonal(H, axis1=1, axis2=2)))
    hd = torch.from_numpy(d if per_matrix else d[0].copy()).to("cuda")
    dS = torch.from_numpy(np.ascontiguousarray(S)).to("cuda")
    with sg.Solver(n, kernel=kernel) as s:
        W, U, info = sg.run_kernel_hdiag(s, hd, dS)
        torch.cuda.synchronize()
    return W.cpu().numpy(), U.cpu().numpy(), info.cpu().numpy()


@pytest.mark.parametrize("kernel", ["unified", "hybrid", "mixed"])
@pytest.mark.parametrize("L", [12, 27, 48])
def test_diagonal_h_entry_point(kernel, L):
    """sunny_geneig_launch_hdiag with one shared Sigma for the batch: the
    split and mixed modes at these orders form M from the diagonal directly
    (H never loaded), the fused mode expands it and must match
    sunny_geneig_launch on the full H bit for bit; all against zhegv with the
    usual bars and the +- pairing."""
    H, S = bogoliubov_pencils(L, 8, 2000 + L)
    W, U, info = _run_hdiag(H, S, kernel)
    assert (info == 0).all()
    scale = np.abs(W).max(axis=1, keepdims=True)
    assert (np.abs(W + W[:, ::-1]) / scale).max() < 1e-12
    Wl, Ul, il = O.zhegv_batch(H, S, threads=8)
    T.assert_parity(H, S, W, U, info, Wl, Ul, tag=f"hdiag {kernel} Sigma/D L={L}")
    if kernel == "unified":
        Wf, Uf, _ = T.run_gpu(H, S, kernel=kernel)
        assert np.array_equal(W, Wf) and np.array_equal(U, Uf)


@pytest.mark.parametrize("kernel,n", [("mixed", 54), ("mixed", 7), ("unified", 54), ("hybrid", 54), ("hybrid", 7),
                                      ("mixed", 96)])
def test_diagonal_h_per_matrix(kernel, n):
    """A different real (indefinite) diagonal per matrix (stride n), including
    a staged order, against zhegv on the full diagonal H."""
    rng = np.random.default_rng(77 + n)
    _, S = O.gen_reference(n, 12, 3300 + n, "random")
    d = rng.uniform(-3.0, 3.0, size=(12, n))
    H = np.zeros((12, n, n), dtype=np.complex128)
    for b in range(12):
        H[b] = np.diag(d[b])
    W, U, info = _run_hdiag(H, S, kernel, per_matrix=True)
    Wl, Ul, il = O.zhegv_batch(H, S, threads=8)
    T.assert_parity(H, S, W, U, info, Wl, Ul, tag=f"hdiag per-matrix {kernel} n={n}")


def test_diagonal_h_rejects_bad_arguments():
    import torch
    import sunny_gen_eigensolver_amd as sg
    with sg.Solver(12, kernel="mixed") as s:
        S = torch.zeros((2, 12, 12), dtype=torch.complex128, device="cuda")
        with pytest.raises(ValueError):
            sg.run_kernel_hdiag(s, torch.zeros((5,), dtype=torch.float64, device="cuda"), S)
        L = sg.lib()
        st = L.sunny_geneig_launch_hdiag(s.handle, 1, -1, 1, 1, 1, 1, 2, 0)
        assert st == sg.STATUS_INVALID_ARG
        st = L.sunny_geneig_launch_hdiag(s.handle, 0, 0, 1, 1, 1, 1, 2, 0)
        assert st == sg.STATUS_INVALID_ARG and "NULL" in sg.last_error()
