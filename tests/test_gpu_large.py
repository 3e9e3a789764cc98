"""GPU parity of the staged kernels (orders 65..192: Cholesky and L^-1 by
panels, M = W H W^H and the back-transform on the matrix cores, the
tridiagonalisation with its triangle in LDS, QL with Z rows in LDS or global
memory, V = Q Z), in the three modes, through the C ABI.

The reference's fixtures at N = 96 and 128 (tests/data/n96_*, n128_*, used by
tests/unified/test_multiN.cpp) are checked against zhegv and their stored
zhegvd eigenvalues with the reference's bars; the benchmark generator covers
N = 65 (partial tiles, two rows per lane), 150 and 192 (FP64: the first
columns of the tridiagonalisation reduced in global memory, Z rows partly in
global memory) with the residual and S-orthogonality bars (<= 50 N eps).
zhegv's INFO for a non-positive-definite S is checked at pivots in the first,
a middle and the last Cholesky panel."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as T

pytestmark = pytest.mark.gpu

MODES = ["unified", "hybrid", "mixed"]
LARGE_FIXTURES = [f"n{n}_b8_s42_{m}" for n in (96, 128) for m in ("random", "nearid")]


@pytest.mark.parametrize("kernel", MODES)
@pytest.mark.parametrize("name", LARGE_FIXTURES)
def test_fixture_vs_lapack(name, kernel):
    meta, H, S = T.fixture_inputs(name)
    W, U, info = T.run_gpu(H, S, kernel=kernel)
    Wl, Ul, il = O.zhegv_batch(H, S, threads=8)
    assert (il == 0).all() and (info == 0).all()
    T.assert_parity(H, S, W, U, info, Wl, Ul, tag=f"{kernel} {name} vs zhegv")
    Wg = T.golden_arrays()[f"{name}__W"]
    rel = np.abs(W - Wg) / np.maximum(np.abs(Wg), 1.0)
    assert rel.max() < T.TOL_EIG


@pytest.mark.parametrize("kernel", MODES)
@pytest.mark.parametrize("n", [65, 150, 192])
def test_baseline_generator(kernel, n):
    H, S = O.gen_spectrum(n, 6, 0x5EED0004 + n)
    Wl, Ul, il = O.zhegv_batch(H, S, threads=8)
    W, U, info = T.run_gpu(H, S, kernel=kernel)
    assert (info == 0).all()
    T.assert_parity(H, S, W, U, info, Wl, Ul, tag=f"{kernel} baseline-gen N={n}")


@pytest.mark.parametrize("kernel", ["unified", "mixed"])
@pytest.mark.parametrize("n", [96, 150])
def test_non_hpd_info_codes(kernel, n):
    """INFO = n + k for the first non-positive leading minor k (1-based), W NaN
    and U zero for that matrix; the rest of the batch is solved normally."""
    H, S = O.gen_reference(n, 6, 700 + n, "random")
    S = S.copy()
    bad = {1: 0, 3: 40, 4: n - 1}
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
