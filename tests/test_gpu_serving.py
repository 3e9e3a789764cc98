"""GPU tests of the serving session (sunny_geneig_session_*): continuous
batching over resident solvers.

- Each request's result is the same alone and in the mix: a set of requests
  of different orders and sizes is solved one at a time, then submitted all
  at once so they overlap on the device; outputs are compared bit for bit
  (every pencil is solved by the same kernels whatever runs beside it) and
  against LAPACK zhegv with the usual bars.
- Host-buffer requests (staged through the session's device arena) give the
  same bits as device-resident ones.
- Admission by HBM footprint: the arena's high-water mark never exceeds the
  budget, and a budget that holds one request at a time serialises them.
- Requests the session cannot serve are rejected with INVALID_ARG."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as T

pytestmark = pytest.mark.gpu

REQS = [(54, 40), (24, 7), (54, 3), (24, 130), (54, 96), (24, 1), (54, 17), (24, 64)]


def _inputs(seed=300):
    out = []
    for k, (n, c) in enumerate(REQS):
        H, S = O.gen_reference(n, c, seed + k, "random")
        out.append((H, S))
    return out


def _dev(arrs):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a)).to("cuda") for a in arrs]


def _outputs(n, c, device):
    import torch
    kw = dict(device=device)
    if device == "cpu":
        kw["pin_memory"] = True
    return (torch.empty((c, n), dtype=torch.float64, **kw),
            torch.empty((c, n, n), dtype=torch.complex128, **kw),
            torch.zeros((c,), dtype=torch.int32, **kw))


@pytest.mark.parametrize("kernel,policy,slice_", [("mixed", "fifo", 0), ("unified", "fifo", 0),
                                                  ("mixed", "srpt", 32), ("unified", "fifo", 16)])
def test_alone_equals_mixed(kernel, policy, slice_):
    """Slices (a request cut into launches of at most slice_ pencils, spread
    over the slots) and either admission order give the same bits as whole
    requests solved alone."""
    import sunny_gen_eigensolver_amd as sg
    data = _inputs()
    with sg.ServingSession([24, 54], kernel=kernel, max_count=256, slots=4, hbm_budget=64 << 20,
                           policy=policy, slice=slice_) as ss:
        alone = []
        for (H, S), (n, c) in zip(data, REQS):
            dH, dS = _dev((H, S))
            W, U, info = _outputs(n, c, "cuda")
            ss.wait(ss.submit(dH, dS, W, U, info))
            alone.append((W.cpu().numpy(), U.cpu().numpy(), info.cpu().numpy()))
        ids, outs = [], []
        for (H, S), (n, c) in zip(data, REQS):
            dH, dS = _dev((H, S))
            W, U, info = _outputs(n, c, "cuda")
            ids.append(ss.submit(dH, dS, W, U, info))
            outs.append((W, U, info))
        lat = [ss.wait(i)[0] for i in ids]
        st = ss.stats()
    assert st["completed"] == 2 * len(REQS)
    assert all(x >= 0 for x in lat)
    for k, ((W, U, info), (Wa, Ua, ia)) in enumerate(zip(outs, alone)):
        assert np.array_equal(W.cpu().numpy(), Wa), f"request {k}: eigenvalues differ alone vs mixed"
        assert np.array_equal(U.cpu().numpy(), Ua), f"request {k}: eigenvectors differ alone vs mixed"
        assert np.array_equal(info.cpu().numpy(), ia)
    for k in (0, 3, 4):
        H, S = data[k]
        Wl, Ul, il = O.zhegv_batch(H, S, threads=8)
        T.assert_parity(H, S, alone[k][0], alone[k][1], alone[k][2], Wl, Ul,
                        tag=f"serving {kernel} request {k}")


@pytest.mark.parametrize("slice_", [0, 24])
def test_host_buffers_match_device_buffers(slice_):
    import torch
    import sunny_gen_eigensolver_amd as sg
    data = _inputs(400)
    with sg.ServingSession([24, 54], kernel="mixed", max_count=256, slots=3, hbm_budget=64 << 20,
                           policy="srpt", slice=slice_) as ss:
        dev_out, host_out, ids = [], [], []
        for (H, S), (n, c) in zip(data, REQS):
            dH, dS = _dev((H, S))
            W, U, info = _outputs(n, c, "cuda")
            ss.wait(ss.submit(dH, dS, W, U, info))
            dev_out.append((W.cpu().numpy(), U.cpu().numpy()))
            hH = torch.from_numpy(np.ascontiguousarray(H)).pin_memory()
            hS = torch.from_numpy(np.ascontiguousarray(S)).pin_memory()
            Wh, Uh, ih = _outputs(n, c, "cpu")
            ids.append(ss.submit(hH, hS, Wh, Uh, ih))
            host_out.append((Wh, Uh, ih, hH, hS))
        for i in ids:
            ss.wait(i)
    for (Wd, Ud), (Wh, Uh, ih, _, _) in zip(dev_out, host_out):
        assert (ih.numpy() == 0).all()
        assert np.array_equal(Wh.numpy(), Wd) and np.array_equal(Uh.numpy(), Ud)


def test_admission_respects_hbm_budget():
    import torch
    import sunny_gen_eigensolver_amd as sg
    n, c = 54, 128
    one = sg.ServingSession.request_bytes(n, c)
    H, S = O.gen_reference(n, c, 77, "random")
    hH = torch.from_numpy(np.ascontiguousarray(H)).pin_memory()
    hS = torch.from_numpy(np.ascontiguousarray(S)).pin_memory()
    for budget, serial in ((one + one // 2, True), (8 * one, False)):
        with sg.ServingSession([n], kernel="unified", max_count=c, slots=4, hbm_budget=budget) as ss:
            keep, ids = [], []
            for _ in range(8):
                W, U, info = _outputs(n, c, "cpu")
                keep.append((W, U, info))
                ids.append(ss.submit(hH, hS, W, U, info))
            for i in ids:
                ss.wait(i)
            st = ss.stats()
        assert st["hbm_peak_bytes"] <= budget
        if serial:
            assert st["max_concurrent"] == 1 and st["hbm_peak_bytes"] == one
        for W, U, info in keep:
            assert np.array_equal(W.numpy(), keep[0][0].numpy())


def test_rejects_what_it_cannot_serve():
    import torch
    import sunny_gen_eigensolver_amd as sg
    with sg.ServingSession([24], kernel="unified", max_count=8, slots=1, hbm_budget=1 << 16) as ss:
        x = torch.zeros((2, 30, 30), dtype=torch.complex128, device="cuda")
        W, U, info = _outputs(30, 2, "cuda")
        with pytest.raises(sg.GeneigError) as e1:
            ss.submit(x, x, W, U, info)                  # order not served
        assert e1.value.status == sg.STATUS_INVALID_ARG
        y = torch.zeros((9, 24, 24), dtype=torch.complex128, device="cuda")
        W, U, info = _outputs(24, 9, "cuda")
        with pytest.raises(sg.GeneigError) as e2:
            ss.submit(y, y, W, U, info)                  # count above max_count
        assert e2.value.status == sg.STATUS_INVALID_ARG
        h = torch.zeros((8, 24, 24), dtype=torch.complex128).pin_memory()
        W, U, info = _outputs(24, 8, "cpu")
        with pytest.raises(sg.GeneigError) as e3:
            ss.submit(h, h, W, U, info)                  # footprint above the budget
        assert e3.value.status == sg.STATUS_INVALID_ARG


def test_staged_orders_in_the_mix():
    """Orders above 64 (the staged kernels) served beside an LDS-resident one:
    bit-identical alone and in the mix, and zhegv parity."""
    import sunny_gen_eigensolver_amd as sg
    reqs = [(96, 9), (24, 5), (150, 3), (96, 2), (150, 1)]
    data = [O.gen_reference(n, c, 500 + k, "random") for k, (n, c) in enumerate(reqs)]
    with sg.ServingSession([24, 96, 150], kernel="mixed", max_count=16, slots=3, hbm_budget=64 << 20) as ss:
        alone = []
        for (H, S), (n, c) in zip(data, reqs):
            dH, dS = _dev((H, S))
            W, U, info = _outputs(n, c, "cuda")
            ss.wait(ss.submit(dH, dS, W, U, info))
            alone.append((W.cpu().numpy(), U.cpu().numpy(), info.cpu().numpy()))
        ids, outs = [], []
        for (H, S), (n, c) in zip(data, reqs):
            dH, dS = _dev((H, S))
            W, U, info = _outputs(n, c, "cuda")
            ids.append(ss.submit(dH, dS, W, U, info))
            outs.append((W, U, info))
        for i in ids:
            ss.wait(i)
    for k, ((W, U, info), (Wa, Ua, ia)) in enumerate(zip(outs, alone)):
        assert np.array_equal(W.cpu().numpy(), Wa) and np.array_equal(U.cpu().numpy(), Ua), k
        assert (ia == 0).all()
    for k in (0, 2):
        H, S = data[k]
        Wl, Ul, il = O.zhegv_batch(H, S, threads=8)
        T.assert_parity(H, S, alone[k][0], alone[k][1], alone[k][2], Wl, Ul, tag=f"serving staged request {k}")



def test_waits_for_the_callers_stream():
    """Inputs written on the caller's stream just before submission, behind a
    long kernel: the session's launch waits for that stream, so the solve sees
    the new inputs (the buffers first hold a non-HPD S that would fail)."""
    import torch
    import sunny_gen_eigensolver_amd as sg
    n, c = 54, 64
    H, S = O.gen_reference(n, c, 611, "random")
    srcH, srcS = _dev((H, S))
    with sg.ServingSession([n], kernel="mixed", max_count=c, slots=2, hbm_budget=1 << 20) as ss:
        W0, U0, i0 = _outputs(n, c, "cuda")
        ss.wait(ss.submit(srcH, srcS, W0, U0, i0))
        side = torch.cuda.Stream()
        dH = torch.zeros_like(srcH)
        dS = torch.zeros_like(srcS)              # zero S: every Cholesky would fail
        torch.cuda.synchronize()
        with torch.cuda.stream(side):
            a = torch.randn((4096, 4096), device="cuda", dtype=torch.float64)
            for _ in range(4):
                a = a @ a / 4096.0                # keep the side stream busy
            dH.copy_(srcH)
            dS.copy_(srcS)
            W, U, info = _outputs(n, c, "cuda")
            rid = ss.submit(dH, dS, W, U, info)
        ss.wait(rid)
        torch.cuda.synchronize()
        del a
    assert (info.cpu().numpy() == 0).all()
    assert np.array_equal(W.cpu().numpy(), W0.cpu().numpy())
    assert np.array_equal(U.cpu().numpy(), U0.cpu().numpy())


def test_submit_validates_buffers():
    """Mismatched devices, shapes, dtypes or layouts are refused before any
    device work (a host W beside device H would otherwise be written through
    a host pointer by the kernels)."""
    import torch
    import sunny_gen_eigensolver_amd as sg
    n, c = 24, 4
    with sg.ServingSession([n], kernel="unified", max_count=8, slots=1, hbm_budget=1 << 20) as ss:
        H = torch.zeros((c, n, n), dtype=torch.complex128, device="cuda")
        W, U, info = _outputs(n, c, "cuda")
        Wc, _, _ = _outputs(n, c, "cpu")
        bad = [
            (H, H, Wc, U, info),                                        # W on the host
            (H, H, W, U.to(torch.complex64), info),                     # U dtype
            (H, H, W[:, :n - 1], U, info),                              # W shape
            (H, H, W, U, info.to(torch.int64)),                         # info dtype
            (H, H.transpose(1, 2), W, U, info),                         # S not contiguous
            (H, H[:2], W, U, info),                                     # S count
        ]
        for k, args in enumerate(bad):
            with pytest.raises(ValueError):
                ss.submit(*args)
