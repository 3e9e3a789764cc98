"""GPU test of sunny_geneig_launch_noalloc, the launch whose workspace comes
from the library's per-device scratch array instead of an allocation: its
outputs equal sunny_geneig_launch's on the same inputs (bit for bit: the
same kernels run, with QL without the rotation log in both at these batch
sizes), on two calls in a row, the second with a smaller batch (stale
scratch contents from the first call must not leak into the second); plus
the split, fused and staged paths and the size limit."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as T

pytestmark = pytest.mark.gpu
MODES = {"unified": 0, "hybrid": 1, "mixed": 2}


def _solve(L, fn, h, H, S):
    import torch
    b, n = H.shape[0], H.shape[1]
    dH = torch.from_numpy(np.ascontiguousarray(H)).to("cuda")
    dS = torch.from_numpy(np.ascontiguousarray(S)).to("cuda")
    W = torch.full((b, n), np.nan, dtype=torch.float64, device="cuda")
    U = torch.full((b, n, n), np.nan, dtype=torch.complex128, device="cuda")
    info = torch.full((b,), -9, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    st = fn(h, dH.data_ptr(), dS.data_ptr(), W.data_ptr(), U.data_ptr(), info.data_ptr(), b,
            int(stream.cuda_stream))
    torch.cuda.synchronize()
    return st, W.cpu().numpy(), U.cpu().numpy(), info.cpu().numpy()


@pytest.mark.parametrize("mode,n,b1,b2", [("mixed", 54, 640, 96), ("hybrid", 54, 640, 33),
                                          ("unified", 54, 300, 17), ("mixed", 96, 40, 7),
                                          ("hybrid", 12, 900, 5)])
def test_noalloc_matches_launch(mode, n, b1, b2):
    import sunny_gen_eigensolver_amd as sg
    L = sg.lib()
    h = L.sunny_geneig_create_ex(n, 0, 0, sg.TUNING_SUGGESTED, sg.CACHE_AUTO, MODES[mode])
    assert h, sg.last_error()
    try:
        for call, (b, seed) in enumerate(((b1, 0x0A11C0 + n), (b2, 0x0A11C1 + n))):
            H, S = O.gen_spectrum(n, b, seed)
            st, Wr, Ur, ir = _solve(L, L.sunny_geneig_launch, h, H, S)
            assert st == sg.STATUS_OK, sg.last_error()
            st, W, U, info = _solve(L, L.sunny_geneig_launch_noalloc, h, H, S)
            assert st == sg.STATUS_OK, sg.last_error()
            assert np.array_equal(info, ir), f"call {call}: info differs"
            assert np.array_equal(W, Wr), f"{mode} n={n} call {call}: W differs from sunny_geneig_launch"
            assert np.array_equal(U, Ur), f"{mode} n={n} call {call}: U differs from sunny_geneig_launch"
            k = min(b, 16)
            Wl, Ul, _ = O.zhegv_batch(H[:k], S[:k], threads=8)
            T.assert_parity(H[:k], S[:k], W[:k], U[:k], info[:k], Wl, Ul, tag=f"noalloc {mode} n={n} call {call}")
    finally:
        L.sunny_geneig_destroy(h)


def test_noalloc_rejects_batches_beyond_the_scratch():
    import sunny_gen_eigensolver_amd as sg
    L = sg.lib()
    h = L.sunny_geneig_create_ex(54, 0, 0, sg.TUNING_SUGGESTED, sg.CACHE_AUTO, MODES["mixed"])
    assert h, sg.last_error()
    try:
        # Far more than 128 MiB of mixed-mode workspace; rejected before any
        # pointer is touched.
        st = L.sunny_geneig_launch_noalloc(h, 1, 1, 1, 1, 1, 100000, 0)
        assert st == sg.STATUS_INVALID_ARG
        msg = sg.last_error()
        assert "scratch" in msg and "134217728" in msg, msg
        assert L.sunny_geneig_launch_noalloc(h, 0, 0, 0, 0, 0, 0, 0) == sg.STATUS_OK
        assert L.sunny_geneig_launch_noalloc(None, 1, 1, 1, 1, 1, 1, 0) == sg.STATUS_INVALID_ARG
        assert L.sunny_geneig_launch_noalloc(h, 1, 1, 1, 1, 1, -1, 0) == sg.STATUS_INVALID_ARG
    finally:
        L.sunny_geneig_destroy(h)
