"""The no-log QL path, forced from the host.

The split and mixed modes run QL in two steps: a recorder logs every sweep's
window and rotations, then the rows of Z are rotated from the log. A matrix
whose QL needs more sweeps than the log holds is solved without a log instead
(geneig_ql_full for orders up to 64, the staged QL kernel's inline path
above). SUNNY_GENEIG_QL_SWEEP_CAP, read when a solver is created, lowers the
log's sweep capacity: cap 0 sends every matrix down that path, a cap of about
the typical sweep count splits a batch between the two. Results are checked
against zhegv with the usual bars."""
import pytest

from oracle import oracle as O
from tests import helpers as T

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kernel", ["hybrid", "mixed"])
@pytest.mark.parametrize("name", ["n54_b8_s42_random", "n32_b8_s42_nearid"])
@pytest.mark.parametrize("cap_per_n", [0.0, 1.2])
def test_lds_orders_no_log_path(monkeypatch, kernel, name, cap_per_n):
    meta, H, S = T.fixture_inputs(name)
    n = H.shape[1]
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
