"""The two QL paths of the split and mixed modes, each forced from the host.

Large batches run QL in two steps: a recorder logs every sweep's window and
rotations, then the rows of Z are rotated from the log. Batches below
SUNNY_GENEIG_QL_LOG_MIN_BATCH (45056 by default) run the no-log kernel
(geneig_ql_full) for every matrix, and so do the matrices whose QL needs more
sweeps than the log holds (for orders above 64, the staged QL kernel's inline
path). Both variables are read when a solver is created: here the minimum
batch is set to 0 so that the small test batches take the log, and
SUNNY_GENEIG_QL_SWEEP_CAP lowers the log's sweep capacity: cap 0 sends every
matrix down the no-log path, a cap of about the typical sweep count splits a
batch between the two. Results are checked against zhegv with the usual
bars."""
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
