"""The two QL paths of the split and mixed modes, each forced from the host.

Large batches run QL in two steps: a recorder logs every sweep's window and
rotations, then the rows of Z are rotated from the log. Batches below
SUNNY_GENEIG_QL_LOG_MIN_BATCH (6144 in the split mode, 12288 in the mixed one,
by default) run the no-log kernel
(geneig_ql_full) for every matrix, and so do the matrices whose QL needs more
sweeps than the log holds (for orders above 64, the staged QL kernel's inline
path). Both variables are read when a solver is created: here the minimum
batch is set to 0 so that the small test batches take the log, and
SUNNY_GENEIG_QL_SWEEP_CAP lowers the log's sweep capacity: cap 0 sends every
matrix down the no-log path, a cap of about the typical sweep count splits a
batch between the two. Results are checked against zhegv with the usual
bars.

At the staged orders up to 128 the recurrence runs one wave per matrix below
SUNNY_GENEIG_QL_LANES_MIN_BATCH matrices per launch and 16 matrices per wave
(one per lane) from there; here the minimum is set to 0 to force the lane
form on small batches (order 150 keeps the one-wave form), with partial last
waves, non-HPD pencils beside HPD ones and the sweep cap splitting the batch
between the log and the inline path."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as T

pytestmark = pytest.mark.gpu

LDS_FIXTURES = [f"n{n}_b8_s42_{m}" for n in (16, 32, 48, 54, 64) for m in ("random", "nearid")]


@pytest.mark.parametrize("kernel", ["hybrid", "mixed"])
@pytest.mark.parametrize("name", LDS_FIXTURES)
def test_logged_path_on_fixtures(monkeypatch, kernel, name):
    monkeypatch.setenv("SUNNY_GENEIG_QL_LOG_MIN_BATCH", "0")
    meta, H, S = T.fixture_inputs(name)
    W, U, info = T.run_gpu(H, S, kernel=kernel)
    Wl, Ul, il = O.zhegv_batch(H, S, threads=8)
    T.assert_parity(H, S, W, U, info, Wl, Ul, tag=f"{kernel} {name} QL through the log")


@pytest.mark.parametrize("kernel", ["hybrid", "mixed"])
@pytest.mark.parametrize("name", ["n54_b8_s42_random", "n32_b8_s42_nearid"])
@pytest.mark.parametrize("cap_per_n", [0.0, 1.2])
def test_lds_orders_no_log_path(monkeypatch, kernel, name, cap_per_n):
    meta, H, S = T.fixture_inputs(name)
    n = H.shape[1]
    monkeypatch.setenv("SUNNY_GENEIG_QL_LOG_MIN_BATCH", "0")
    monkeypatch.setenv("SUNNY_GENEIG_QL_SWEEP_CAP", str(int(cap_per_n * n)))
    W, U, info = T.run_gpu(H, S, kernel=kernel)
    Wl, Ul, il = O.zhegv_batch(H, S, threads=8)
    T.assert_parity(H, S, W, U, info, Wl, Ul, tag=f"{kernel} {name} sweep cap {cap_per_n} N")


@pytest.mark.parametrize("kernel", ["unified", "mixed"])
@pytest.mark.parametrize("cap_per_n", [0.0, 1.2])
def test_staged_order_no_log_path(monkeypatch, kernel, cap_per_n):
    meta, H, S = T.fixture_inputs("n96_b8_s42_random")
    n = H.shape[1]
    monkeypatch.setenv("SUNNY_GENEIG_QL_SWEEP_CAP", str(int(cap_per_n * n)))
    W, U, info = T.run_gpu(H, S, kernel=kernel)
    Wl, Ul, il = O.zhegv_batch(H, S, threads=8)
    T.assert_parity(H, S, W, U, info, Wl, Ul, tag=f"{kernel} n96 sweep cap {cap_per_n} N")


@pytest.mark.parametrize("kernel", ["unified", "mixed"])
@pytest.mark.parametrize("n", [65, 96, 128, 150])
def test_staged_lane_recorder(monkeypatch, kernel, n):
    monkeypatch.setenv("SUNNY_GENEIG_QL_LANES_MIN_BATCH", "0")
    H, S = O.gen_spectrum(n, 70, 0x1A7E0000 + n)
    W, U, info = T.run_gpu(H, S, kernel=kernel)
    Wl, Ul, il = O.zhegv_batch(H, S, threads=8)
    assert (info == 0).all()
    T.assert_parity(H, S, W, U, info, Wl, Ul, tag=f"{kernel} n={n} lane recorder")


@pytest.mark.parametrize("kernel", ["unified", "mixed"])
def test_staged_lane_recorder_non_hpd_and_cap(monkeypatch, kernel):
    """Lanes whose Cholesky failed sit out (INFO = n + k, W NaN, U zero); a
    sweep cap near the typical sweep count sends part of the batch to the
    inline path."""
    n = 96
    monkeypatch.setenv("SUNNY_GENEIG_QL_LANES_MIN_BATCH", "0")
    monkeypatch.setenv("SUNNY_GENEIG_QL_SWEEP_CAP", str(int(1.2 * n)))
    H, S = O.gen_reference(n, 20, 900 + n, "random")
    S = S.copy()
    bad = {2: 0, 7: 50, 19: n - 1}
    for b, k in bad.items():
        Sm = np.eye(n, dtype=np.complex128)
        Sm[k, k] = -1.0
        S[b] = Sm.T
    Wl, Ul, il = O.zhegv_batch(H, S, threads=8)
    W, U, info = T.run_gpu(H, S, kernel=kernel)
    for b, k in bad.items():
        assert int(info[b]) == n + k + 1 == int(il[b]), (b, info[b], il[b])
        assert np.isnan(W[b]).all() and (U[b] == 0).all()
    good = [b for b in range(20) if b not in bad]
    T.assert_parity(H[good], S[good], W[good], U[good], info[good], Wl[good], Ul[good],
                    tag=f"{kernel} n={n} lane recorder, sweep cap, beside non-HPD")
