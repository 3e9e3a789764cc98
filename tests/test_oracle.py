"""Pins the CPU oracle before it is trusted as the GPU checker.

- generators: bit-exact against the reference's fixtures (SHA-256 of H, S
  stored in tests/golden/fixtures.json by make_golden.py);
- LAPACK binding: reproduces the reference's stored zhegvd eigenvalues;
- restatement of the GPU algorithm, stage by stage: potrf vs zpotrf, the
  two-sided solve vs zhegst, tridiagonalisation by invariants (unitary Q,
  Q^H M Q tridiagonal and equal to (d, e), eigenvalues preserved), QL vs dsteqr,
  closed forms on tiny inputs (2x2, diagonal pencils, S = I), end to end vs
  zhegv and the golden eigenvectors;
- metrics: planted mistakes in the outputs are caught.
"""
import hashlib

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as T

EPS = 2.0 ** -52


def test_generator_bit_exact_all_fixtures():
    for name, meta in T.fixtures_index().items():
        if meta["n"] > 64 and meta["s_mode"] != "random":
            continue  # one large near-identity case is enough (kept below)
        _, H, S = T.fixture_inputs(name)
        assert hashlib.sha256(H.tobytes()).hexdigest() == meta["sha256_H"], name
        assert hashlib.sha256(S.tobytes()).hexdigest() == meta["sha256_S"], name


def test_generator_bit_exact_n96_nearid():
    _, H, S = T.fixture_inputs("n96_b8_s42_nearid")
    meta = T.fixtures_index()["n96_b8_s42_nearid"]
    assert hashlib.sha256(S.tobytes()).hexdigest() == meta["sha256_S"]


@pytest.mark.parametrize("name", ["n16_b8_s42_random", "n54_b8_s42_nearid",
                                  "n54_b32_s123_random", "n64_b8_s42_random"])
def test_lapack_reproduces_golden_eigenvalues(name):
    _, H, S = T.fixture_inputs(name)
    W, U, info = O.zhegv_batch(H, S, threads=8)
    assert (info == 0).all()
    Wg = T.golden_arrays()[f"{name}__W"]
    assert (np.abs(W - Wg) / np.maximum(np.abs(Wg), 1)).max() < 1e-13


@pytest.mark.parametrize("name", ["n16_b8_s42_random", "n54_b8_s42_random"])
def test_restatement_vs_golden_vectors(name):
    _, H, S = T.fixture_inputs(name)
    g = T.golden_arrays()
    W, U, info = O.rs_zhegv_batch(H, S, threads=8)
    T.assert_parity(H, S, W, U, info, g[f"{name}__W"], g[f"{name}__U"], tag=f"restate {name}")


@pytest.mark.parametrize("n", [1, 2, 3, 16, 33, 54, 64])
def test_restatement_vs_zhegv(n):
    H, S = O.gen_reference(n, 6, 1000 + n, "random")
    W, U, info = O.rs_zhegv_batch(H, S)
    Wl, Ul, il = O.zhegv_batch(H, S)
    assert (il == 0).all()
    T.assert_parity(H, S, W, U, info, Wl, Ul, tag=f"restate n={n}")


def test_restatement_spectrum_generator():
    H, S = O.gen_spectrum(54, 8, 0x5EED0001)
    W, U, info = O.rs_zhegv_batch(H, S)
    Wl, Ul, il = O.zhegv_batch(H, S)
    T.assert_parity(H, S, W, U, info, Wl, Ul, tag="restate spectrum")


def _mat(a):
    """Column-major storage -> numpy matrix."""
    return np.asarray(a).T


def test_potrf_vs_zpotrf():
    """tests/static/test_chol_n54.cpp:141-162: max_abs / ||L||_F < 1e-12."""
    _, H, S = T.fixture_inputs("n54_b8_s42_random")
    for b in range(8):
        L1, i1 = O.rs_potrf(S[b])
        L2, i2 = O.zpotrf(S[b])
        assert i1 == 0 and i2 == 0
        assert np.abs(L1 - L2).max() / np.linalg.norm(L2) < 1e-12
        assert np.abs(np.triu(_mat(L1), 1)).max() == 0.0


def test_reduce_vs_zhegst():
    """tests/static/test_reduce_n54.cpp:134-149: lower triangle < 1e-11."""
    _, H, S = T.fixture_inputs("n54_b8_s42_random")
    for b in range(8):
        L, _ = O.zpotrf(S[b])
        M1 = O.rs_reduce(L, H[b])
        M2, info = O.zhegst(H[b], L)
        assert info == 0
        lo1, lo2 = np.tril(_mat(M1)), np.tril(_mat(M2))
        assert np.abs(lo1 - lo2).max() / np.abs(lo2).max() < 1e-11


def test_hetrd_invariants():
    """Q unitary, Q^H M Q = T (tridiagonal with the returned d, e)."""
    _, H, S = T.fixture_inputs("n54_b8_s42_random")
    L, _ = O.rs_potrf(S[0])
    M = O.rs_reduce(L, H[0])
    A, d, e, tau = O.rs_hetrd(M)
    n = 54
    Q = O.rs_apply_q(A, tau, np.eye(n, dtype=np.complex128).T.copy())
    Qm = _mat(Q)
    assert np.abs(Qm.conj().T @ Qm - np.eye(n)).max() < 50 * n * EPS
    Mm = _mat(M)
    Tm = Qm.conj().T @ Mm @ Qm
    Tref = np.diag(d) + np.diag(e[:-1], -1) + np.diag(e[:-1], 1)
    assert np.abs(Tm - Tref).max() < 50 * n * EPS * np.abs(Mm).max()
    wl = np.linalg.eigvalsh(0.5 * (Mm + Mm.conj().T))
    wt = np.linalg.eigvalsh(Tref)
    assert np.abs(np.sort(wt) - wl).max() < 1e-12 * np.abs(wl).max()


def test_tql_vs_dsteqr():
    rng = np.random.default_rng(3)
    for n in (1, 2, 5, 54, 64):
        d = rng.standard_normal(n)
        e = rng.standard_normal(n)
        w1, Z1, i1 = O.rs_tql(d, e)
        w2, Z2, i2 = O.dsteqr(d, e)
        assert i1 == 0 and i2 == 0
        assert np.abs(w1 - w2).max() < 1e-13 * max(1, np.abs(w2).max())
        Zm = _mat(Z1)
        assert np.abs(Zm.T @ Zm - np.eye(n)).max() < 50 * max(n, 1) * EPS
        # Z diagonalises T.
        Tm = np.diag(d) + np.diag(e[: n - 1], -1) + np.diag(e[: n - 1], 1)
        assert np.abs(Tm @ Zm - Zm * w1).max() < 1e-12 * max(1, np.abs(w1).max())


def test_tql_clustered_and_zero_offdiagonal():
    # Glued Wilkinson-like matrix with tiny couplings and an exact split.
    n = 21
    d = np.abs(np.arange(n) - 10).astype(float)
    e = np.ones(n)
    e[7] = 0.0
    e[14] = 1e-14
    w1, Z1, i1 = O.rs_tql(d, e)
    w2, _, _ = O.dsteqr(d, e)
    assert i1 == 0
    assert np.abs(w1 - w2).max() < 1e-13 * np.abs(w2).max()


def test_two_by_two_closed_form():
    """H = [[a, c], [conj(c), b]], S = I: eigenvalues (a+b)/2 -+ sqrt(((a-b)/2)^2 + |c|^2)."""
    a, b, c = 1.5, -0.25, 0.3 - 0.7j
    Hm = np.array([[a, c], [np.conj(c), b]])
    H = Hm.T.reshape(1, 2, 2).copy()
    S = np.eye(2, dtype=np.complex128).reshape(1, 2, 2).copy()
    W, U, info = O.rs_zhegv_batch(H, S)
    r = np.sqrt(((a - b) / 2) ** 2 + abs(c) ** 2)
    assert np.allclose(W[0], [(a + b) / 2 - r, (a + b) / 2 + r], rtol=0, atol=1e-15)


def test_diagonal_pencil():
    """Diagonal H and S: lambda = h_ii / s_ii, U = diag(1/sqrt(s_ii)) up to order/phase."""
    n = 9
    h = np.linspace(-3, 5, n)
    s = np.linspace(0.5, 4, n)
    H = np.diag(h).astype(np.complex128).reshape(1, n, n).copy()
    S = np.diag(s).astype(np.complex128).reshape(1, n, n).copy()
    W, U, info = O.rs_zhegv_batch(H, S)
    assert info[0] == 0
    assert np.allclose(W[0], np.sort(h / s), atol=1e-14)
    Um = _mat(U[0])
    assert np.allclose(np.abs(Um).max(axis=0), np.sort_complex(np.zeros(n)) + 1 / np.sqrt(s)[np.argsort(h / s)],
                       atol=1e-14)


def test_identity_s_reduces_to_zheev():
    _, H, _ = T.fixture_inputs("n16_b8_s42_random")
    n = 16
    S = np.broadcast_to(np.eye(n, dtype=np.complex128), H.shape).copy()
    W, U, info = O.rs_zhegv_batch(H, S)
    for b in range(H.shape[0]):
        w2, _, i2 = O.zheev(H[b])
        assert np.abs(W[b] - w2).max() < 1e-13 * np.abs(w2).max()


def test_non_hpd_info_codes():
    """zhegv INFO = N + k for the first leading minor k of S that is not
    positive definite (documented LAPACK convention)."""
    n = 12
    H, _ = O.gen_reference(n, 1, 5, "random")
    for k in (0, 4, 11):
        S = np.eye(n, dtype=np.complex128)
        S[k, k] = -1.0
        S = S.reshape(1, n, n).copy()
        _, _, il = O.zhegv_batch(H, S)
        _, _, ir = O.rs_zhegv_batch(H, S)
        assert il[0] == n + k + 1
        assert ir[0] == n + k + 1
    # Rank-deficient S = B B^H with B n x (n-3): first failing minor is n-2.
    rng = np.random.default_rng(1)
    Bm = rng.standard_normal((n, n - 3)) + 1j * rng.standard_normal((n, n - 3))
    S = (Bm @ Bm.conj().T).T.reshape(1, n, n).copy()
    _, _, il = O.zhegv_batch(H, S)
    _, _, ir = O.rs_zhegv_batch(H, S)
    assert il[0] > n and ir[0] > n


def test_metrics_catch_planted_mistakes():
    _, H, S = T.fixture_inputs("n16_b8_s42_random")
    W, U, info = O.zhegv_batch(H, S)
    good = T.metrics(H, S, W, U, W, U)
    assert good[:, 0].max() == 0 and good[:, 1].max() < 1e-13
    # Wrong eigenvalue.
    W2 = W.copy()
    W2[0, 3] += 1e-8
    assert T.metrics(H, S, W2, U, W, U)[0, 0] > 1e-9
    assert T.metrics(H, S, W2, U, W, U)[0, 3] > 1e-12
    # Swapped eigenvector columns.
    U2 = U.copy()
    U2[0, [2, 3]] = U2[0, [3, 2]]
    m = T.metrics(H, S, W, U2, W, U)
    assert m[0, 1] > 0.5 and m[0, 3] > 1e-3
    # Loss of S-normalisation.
    U3 = U.copy() * 1.001
    assert T.metrics(H, S, W, U3, W, U)[0, 4] > 1e-3
    # Transposed instead of conjugate-transposed factor: residual blows up.
    U4 = np.conj(U)
    assert T.metrics(H, S, W, U4, W, U)[:, 3].max() > 1e-6


def _spectrum_properties(H, S, n, cond_s, tag):
    """Checks of the benchmark generator's stated law on one batch (matrices
    stored column-major: the stored array is the transpose)."""
    npairs = min(6, n // 2)
    for b in range(H.shape[0]):
        h, s = H[b].T, S[b].T
        assert np.abs(h - h.conj().T).max() <= 1e-14 * np.abs(h).max(), f"{tag}: H not Hermitian"
        assert np.abs(s - s.conj().T).max() <= 1e-14 * np.abs(s).max(), f"{tag}: S not Hermitian"
        lam = np.linalg.eigvalsh(h)
        assert lam.min() >= -10.0 - 1e-12 and lam.max() <= 10.0 + 1e-12, f"{tag}: lambda outside [-10, 10]"
        gaps = np.diff(lam)
        near = np.abs(gaps - 1e-8) <= 1e-12
        assert near.sum() >= npairs, f"{tag}: {near.sum()} pairs split by 1e-8, expected {npairs}"
        assert (gaps[~near] > 1e-9).all() or npairs == 0, f"{tag}: unexpected near-degenerate pair"
        sig = np.linalg.eigvalsh(s)
        assert sig.min() > 0, f"{tag}: S not positive definite"
        cond = sig.max() / sig.min()
        if cond_s:
            assert abs(cond / cond_s - 1.0) <= 1e-8, f"{tag}: cond(S) {cond:.6e} != {cond_s:.6e}"
            want = cond_s ** (np.arange(n) / max(n - 1, 1))
            assert np.allclose(sig, want, rtol=1e-10), f"{tag}: S spectrum not geometric"
        else:
            # The stated band (1e2..1e3) is for the benchmark's order, 54; the
            # ratio of B B^H's extreme eigenvalues shrinks at small orders.
            lo = 1e2 if n >= 32 else 1e1
            assert lo <= cond <= 1e3, f"{tag}: cond(S) {cond:.3e} outside [{lo:g}, 1e3]"


@pytest.mark.parametrize("n", [12, 54, 96])
def test_spectrum_generator_law(n):
    """The benchmark generator (oracle restatement used by the tests): H = Q
    diag(lambda) Q^H with lambda in [-10, 10] and 6 pairs split by 1e-8;
    S = B B^H + 0.05 N I with cond(S) in [1e2, 1e3]; with cond_s, S's spectrum
    is geometric from 1 to cond_s exactly."""
    H, S = O.gen_spectrum(n, 6, 0x5EED0000 + n)
    _spectrum_properties(H, S, n, 0.0, f"oracle n={n}")
    for c in (1e1, 1e4):
        H, S = O.gen_spectrum(n, 3, 0x5EED1000 + n, cond_s=c)
        _spectrum_properties(H, S, n, c, f"oracle n={n} cond {c:g}")


@pytest.mark.parametrize("n", [12, 54])
def test_device_spectrum_generator_law(n):
    """The device generator the benchmark times on (torch, run here on the
    CPU device) obeys the same law, including a per-matrix cond(S)."""
    import torch
    from sunny_gen_eigensolver_amd import inputs
    H, S = inputs.spectrum_batch(n, 4, 0x5EED0002, device="cpu")
    _spectrum_properties(H.numpy(), S.numpy(), n, 0.0, f"device n={n}")
    conds = torch.tensor([1e1, 3e2, 1e4], dtype=torch.float64)
    H, S = inputs.spectrum_batch(n, 3, 0x5EED0003, device="cpu", cond_s=conds)
    for b in range(3):
        _spectrum_properties(H.numpy()[b:b + 1], S.numpy()[b:b + 1], n, float(conds[b]),
                             f"device n={n} cond {float(conds[b]):g}")


def test_spectrum_property_checks_catch_planted_mistakes():
    """A dropped pair split, a lambda range off by a factor and a wrong S
    spectrum each fail the checks above (planted in the oracle's output)."""
    n = 24
    H, S = O.gen_spectrum(n, 1, 0x5EED2000)
    Q = np.linalg.eigh(H[0].T)[1]
    lam = np.linalg.eigvalsh(H[0].T)
    bad = [lam.copy(), lam * 1.5]
    i = int(np.argmin(np.abs(np.diff(lam) - 1e-8)))  # one of the split pairs
    bad[0][i + 1] = bad[0][i] + 1e-3                 # that pair no longer split by 1e-8
    for lb in bad:
        Hb = (Q * lb) @ Q.conj().T
        Hb = 0.5 * (Hb + Hb.conj().T)
        with pytest.raises(AssertionError):
            _spectrum_properties(Hb.T[None], S, n, 0.0, "planted H")
    Hc, Sc = O.gen_spectrum(n, 1, 0x5EED2001, cond_s=1e3)
    with pytest.raises(AssertionError):
        _spectrum_properties(Hc, Sc, n, 1e4, "planted cond")
