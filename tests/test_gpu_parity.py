"""GPU parity of the fused (unified), split (hybrid) and mixed-precision
(FP32 stages + FP64 refinement) modes against FP64 LAPACK zhegv, through the
C ABI.

Mirrors the reference's tests: tests/unified/test_n54.cpp, test_multiN.cpp,
tests/static/test_full_batched_n54.cpp (B in {1, 8, 32} on n54_b32_s123 and a
synthetic B=1024 batch with seed 0xC0FFEE0042 against live zhegv),
tests/hybrid/test_hybrid_n54.cpp and test_hybrid_multiN.cpp. Bars: eig_rel
and phase_err < 1e-10 (columns with gap < 1e-6 excluded), as in the
reference, plus residual and S-orthogonality <= 50 N eps (1-norms).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as T

pytestmark = pytest.mark.gpu

LDS_FIXTURES = [f"n{n}_b8_s42_{m}" for n in (16, 32, 48, 54, 64) for m in ("random", "nearid")]


MODES = ["unified", "hybrid", "mixed"]


@pytest.mark.parametrize("kernel", MODES)
@pytest.mark.parametrize("name", LDS_FIXTURES)
def test_fixture_vs_lapack(name, kernel):
    meta, H, S = T.fixture_inputs(name)
    W, U, info = T.run_gpu(H, S, kernel=kernel)
    Wl, Ul, il = O.zhegv_batch(H, S, threads=8)
    assert (il == 0).all()
    T.assert_parity(H, S, W, U, info, Wl, Ul, tag=f"{kernel} {name} vs zhegv")
    # Against the reference's committed zhegvd eigenvalues as well.
    Wg = T.golden_arrays()[f"{name}__W"]
    rel = np.abs(W - Wg) / np.maximum(np.abs(Wg), 1.0)
    assert rel.max() < T.TOL_EIG


@pytest.mark.parametrize("kernel", MODES)
def test_fixture_vectors_vs_golden(kernel):
    """Phase check against the reference's stored eigenvectors."""
    g = T.golden_arrays()
    for name in ("n16_b8_s42_random", "n54_b8_s42_random"):
        meta, H, S = T.fixture_inputs(name)
        W, U, info = T.run_gpu(H, S, kernel=kernel)
        T.assert_parity(H, S, W, U, info, g[f"{name}__W"], g[f"{name}__U"],
                        tag=f"{kernel} {name} vs golden")


@pytest.mark.parametrize("batch", [1, 8, 32])
def test_n54_b32_slices(batch):
    """tests/static/test_full_batched_n54.cpp:356-361."""
    meta, H, S = T.fixture_inputs("n54_b32_s123_random")
    Wl, Ul, _ = O.zhegv_batch(H[:batch], S[:batch], threads=8)
    W, U, info = T.run_gpu(H[:batch], S[:batch])
    T.assert_parity(H[:batch], S[:batch], W, U, info, Wl, Ul, tag=f"B={batch}")


@pytest.mark.parametrize("kernel", MODES)
def test_synthetic_1024(kernel):
    """tests/static/test_full_batched_n54.cpp:222-334: B=1024, seed 0xC0FFEE0042."""
    H, S = O.gen_reference(54, 1024, 0xC0FFEE0042, "random")
    Wl, Ul, il = O.zhegv_batch(H, S, threads=8)
    assert (il == 0).all()
    W, U, info = T.run_gpu(H, S, kernel=kernel)
    T.assert_parity(H, S, W, U, info, Wl, Ul, tag=f"{kernel} B=1024")


@pytest.mark.parametrize("kernel", MODES)
def test_baseline_generator_n54(kernel):
    """The benchmark's input family (Q diag(lambda) Q^H with near-degenerate
    pairs, S = B B^H + 0.05 N I) against zhegv."""
    H, S = O.gen_spectrum(54, 64, 0x5EED0002)
    Wl, Ul, il = O.zhegv_batch(H, S, threads=8)
    W, U, info = T.run_gpu(H, S, kernel=kernel)
    T.assert_parity(H, S, W, U, info, Wl, Ul, tag=f"{kernel} baseline-gen N=54")


@pytest.mark.parametrize("kernel", MODES)
@pytest.mark.parametrize("n", [12, 54])
def test_non_hpd_info_codes(kernel, n):
    """zhegv's INFO for S that is not positive definite: n + k for the first
    non-positive leading minor k (1-based), W = NaN and U = 0 for that matrix;
    the other matrices of the batch are unaffected."""
    H, S = O.gen_reference(n, 6, 900 + n, "random")
    S = S.copy()
    bad = {1: 0, 3: n // 2, 4: n - 1}               # matrix -> failing pivot (0-based)
    for b, k in bad.items():
        Sm = np.eye(n, dtype=np.complex128)
        Sm[k, k] = -1.0
        S[b] = Sm.T
    Wl, Ul, il = O.zhegv_batch(H, S, threads=8)
    W, U, info = T.run_gpu(H, S, kernel=kernel)
    for b, k in bad.items():
        assert int(info[b]) == n + k + 1 == int(il[b]), (b, info[b], il[b])
        assert np.isnan(W[b]).all() and (U[b] == 0).all()
    good = [b for b in range(6) if b not in bad]
    T.assert_parity(H[good], S[good], W[good], U[good], info[good], Wl[good], Ul[good],
                    tag=f"{kernel} beside non-HPD n={n}")


@pytest.mark.parametrize("kernel", MODES)
@pytest.mark.parametrize("n", [1, 2, 3, 5, 9, 13])
def test_tiny_orders_vs_lapack(kernel, n):
    """Orders far below a 16-row tile or a wave (one matrix element, 2 x 2, odd
    orders) in every mode: the padded MFMA tiles, the partial panels and the
    lane masks all see mostly-empty ranges."""
    H, S = O.gen_reference(n, 16, 4200 + n, "random")
    Wl, Ul, il = O.zhegv_batch(H, S, threads=8)
    W, U, info = T.run_gpu(H, S, kernel=kernel)
    assert (il == 0).all()
    T.assert_parity(H, S, W, U, info, Wl, Ul, tag=f"{kernel} tiny n={n}")


@pytest.mark.parametrize("kernel", MODES)
@pytest.mark.parametrize("n", [54, 96])
def test_ill_conditioned_s(kernel, n):
    """cond(S) log-uniform in [1e1, 1e4] (S = Q_S diag(sigma) Q_S^H, sigma
    geometric from 1 to cond) at an LDS-resident and a staged order, every
    mode: the mixed mode's FP32 stages see M = L^-1 H L^-H, whose conditioning
    grows with cond(S), and the refinement must still land within the
    conditioning-aware bars against zhegv on the same pencils."""
    rng = np.random.default_rng(0x1C0 + n)
    b = 24
    conds = np.exp(rng.uniform(np.log(1e1), np.log(1e4), size=b))
    conds[0] = 1e4
    Hs, Ss = [], []
    for k in range(b):
        H1, S1 = O.gen_spectrum(n, 1, 0xBAD5 + 97 * n + k, cond_s=float(conds[k]))
        Hs.append(H1[0])
        Ss.append(S1[0])
    H, S = np.stack(Hs), np.stack(Ss)
    W, U, info = T.run_gpu(H, S, kernel=kernel)
    T.assert_parity_conditioned(H, S, W, U, info, tag=f"{kernel} n={n} cond(S) <= 1e4")


@pytest.mark.parametrize("n", [54, 96])
def test_refinement_convergence_test_reports_through_info(n, monkeypatch):
    """The mixed mode's refinement checks its last correction: with the
    tolerance forced to (almost) zero from the host every matrix reports
    info = n (W and U still the refined iterate, within the bars); at the
    default tolerance none does."""
    H, S = O.gen_spectrum(n, 16, 0xF1A6 + n)
    monkeypatch.setenv("SUNNY_GENEIG_REFINE_TOL", "1e-300")
    W, U, info = T.run_gpu(H, S, kernel="mixed")
    assert (info == n).all(), info
    Wl, Ul, _ = O.zhegv_batch(H, S, threads=8)
    T.assert_parity(H, S, W, U, np.zeros_like(info), Wl, Ul, tag=f"flagged mixed n={n}")
    monkeypatch.delenv("SUNNY_GENEIG_REFINE_TOL")
    W2, U2, info2 = T.run_gpu(H, S, kernel="mixed")
    assert (info2 == 0).all()
    assert np.array_equal(W, W2) and np.array_equal(U, U2)
