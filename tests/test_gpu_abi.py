"""The C ABI on live handles (GPU), mirroring the reference's
tests/common/test_c_api.cpp:124-219, tests/unified/test_streams_n54.cpp and
the code-object cache tests tests/unified/test_cache_n54.cpp:144-195 and
test_cache_invalidation.cpp:238-253.

- create / introspection / launch of the n54_b8_s42_random fixture through
  sunny_geneig_create(_ex) + sunny_geneig_launch in every kernel mode, then
  validation against the fixture's stored eigenpairs (eig_rel and phase
  < 1e-10, as the reference);
- explicit matrices-per-workgroup requests: 1, 2 and 4 are honoured (odd
  batch sizes included), larger ones return NULL with a message;
- sunny_geneig_launch_chunked over 1, 2 and 4 streams: bit-identical to one
  launch of the whole batch, zhegv spot checks, argument errors;
- the cache layers cold (compile) -> warm disk -> in-process with
  bit-identical W and U, and a code object whose ELF magic is overwritten is
  detected, recompiled and rewritten with the same results."""
import ctypes
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as T

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODES = {"unified": 0, "hybrid": 1, "mixed": 2}


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _launch(L, h, H, S):
    """sunny_geneig_launch on the null stream with torch-allocated buffers."""
    import torch
    b, n = H.shape[0], H.shape[1]
    dH, dS = _dev(H), _dev(S)
    W = torch.empty((b, n), dtype=torch.float64, device="cuda")
    U = torch.empty((b, n, n), dtype=torch.complex128, device="cuda")
    info = torch.full((b,), -7, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    st = L.sunny_geneig_launch(h, dH.data_ptr(), dS.data_ptr(), W.data_ptr(), U.data_ptr(),
                               info.data_ptr(), b, 0)
    torch.cuda.synchronize()
    return st, W.cpu().numpy(), U.cpu().numpy(), info.cpu().numpy()


@pytest.mark.parametrize("mode", list(MODES))
def test_live_handle_fixture_launch(mode):
    import sunny_gen_eigensolver_amd as sg
    L = sg.lib()
    name = "n54_b8_s42_random"
    meta, H, S = T.fixture_inputs(name)
    if mode == "unified":
        h = L.sunny_geneig_create(54, 0, 1, sg.TUNING_SUGGESTED, sg.CACHE_AUTO)
    else:
        h = L.sunny_geneig_create_ex(54, 0, 1, sg.TUNING_SUGGESTED, sg.CACHE_AUTO, MODES[mode])
    assert h, sg.last_error()
    try:
        assert L.sunny_geneig_matrix_size(h) == 54
        assert L.sunny_geneig_batches_per_block(h) == 1       # explicit request / split semantics
        assert L.sunny_geneig_block_dim_x(h) > 0
        assert L.sunny_geneig_shared_mem_bytes(h) != 0
        assert L.sunny_geneig_path_name(h).decode() == mode
        assert L.sunny_geneig_kernel_mode(h) == MODES[mode]
        assert L.sunny_geneig_cache_layer(h).decode() in ("in-process", "disk", "compile")
        st, W, U, info = _launch(L, h, H, S)
        assert st == sg.STATUS_OK, sg.last_error()
        assert (info == 0).all()
        g = T.golden_arrays()
        m = T.metrics(H, S, W, U, g[f"{name}__W"], g[f"{name}__U"])
        assert m[:, 0].max() < T.TOL_EIG and m[:, 1].max() < T.TOL_PHASE, m[:, :2].max(axis=0)
    finally:
        L.sunny_geneig_destroy(h)


def test_split_modes_report_one_matrix_per_block():
    """The reference's split handle reports batches_per_block 1 whatever was
    requested (src/common/c_api.cpp:41-49); the mixed mode does the same."""
    import sunny_gen_eigensolver_amd as sg
    for mode in ("hybrid", "mixed"):
        with sg.Solver(54, kernel=mode, bpb_request=4) as s:
            assert s.batches_per_block == 1
        with sg.Solver(24, kernel=mode) as s:
            assert s.batches_per_block == 1


@pytest.mark.parametrize("bpb", [1, 2, 4])
def test_explicit_matrices_per_block(bpb):
    """An explicit request fixes the workgroup (one wave per matrix); a batch
    that is not a multiple of it leaves a partial last workgroup."""
    import sunny_gen_eigensolver_amd as sg
    L = sg.lib()
    h = L.sunny_geneig_create(54, 0, bpb, sg.TUNING_SUGGESTED, sg.CACHE_AUTO)
    assert h, sg.last_error()
    try:
        assert L.sunny_geneig_batches_per_block(h) == bpb
        assert L.sunny_geneig_block_dim_x(h) == 64 * bpb
        H, S = O.gen_reference(54, 7, 8100 + bpb, "random")
        st, W, U, info = _launch(L, h, H, S)
        assert st == sg.STATUS_OK
        Wl, Ul, _ = O.zhegv_batch(H, S, threads=8)
        T.assert_parity(H, S, W, U, info, Wl, Ul, tag=f"bpb={bpb} B=7")
    finally:
        L.sunny_geneig_destroy(h)


@pytest.mark.parametrize("bpb", [8, 16])
def test_unsupported_matrices_per_block_is_refused(bpb):
    import sunny_gen_eigensolver_amd as sg
    L = sg.lib()
    h = L.sunny_geneig_create(16, 0, bpb, sg.TUNING_SUGGESTED, sg.CACHE_AUTO)
    assert not h
    msg = sg.last_error()
    assert f"batches_per_block_request={bpb}" in msg and "at most 4" in msg, msg


@pytest.mark.parametrize("mode", list(MODES))
def test_chunked_streams_match_one_launch(mode):
    """tests/unified/test_streams_n54.cpp: chunks round-robin over 1, 2 and 4
    streams give exactly the results of one launch over the whole batch."""
    import torch
    import sunny_gen_eigensolver_amd as sg
    n, total, chunk = 54, 1200, 256
    H, S = O.gen_reference(n, total, 0x57EA, "random")
    dH, dS = _dev(H), _dev(S)
    with sg.Solver(n, kernel=mode) as s:
        W0, U0, I0 = sg.run_kernel(s, dH, dS)
        torch.cuda.synchronize()
        W0, U0, I0 = W0.cpu().numpy(), U0.cpu().numpy(), I0.cpu().numpy()
        assert (I0 == 0).all()
        for ns in (1, 2, 4):
            streams = [torch.cuda.Stream() for _ in range(ns)]
            W = torch.full((total, n), np.nan, dtype=torch.float64, device="cuda")
            U = torch.zeros((total, n, n), dtype=torch.complex128, device="cuda")
            info = torch.full((total,), -1, dtype=torch.int32, device="cuda")
            torch.cuda.synchronize()
            sg.run_kernel_chunked(s, dH, dS, W, U, info, chunk, streams)
            torch.cuda.synchronize()
            assert np.array_equal(info.cpu().numpy(), I0), ns
            assert np.array_equal(W.cpu().numpy(), W0), f"{mode}: W differs with {ns} streams"
            assert np.array_equal(U.cpu().numpy(), U0), f"{mode}: U differs with {ns} streams"
    spot = [0, 255, 256, 511, 700, 1023, 1024, 1199]        # chunk edges
    Wl, Ul, _ = O.zhegv_batch(H[spot], S[spot], threads=8)
    T.assert_parity(H[spot], S[spot], W0[spot], U0[spot], I0[spot], Wl, Ul, tag=f"chunked {mode} spots")


def test_chunked_argument_errors():
    import sunny_gen_eigensolver_amd as sg
    L = sg.lib()
    one = (ctypes.c_uint64 * 1)(0)
    assert L.sunny_geneig_launch_chunked(None, 1, 1, 1, 1, 1, 8, 4, one, 1) == sg.STATUS_INVALID_ARG
    with sg.Solver(16) as s:
        for total, chunk, streams, ns in ((8, 0, one, 1), (-1, 4, one, 1), (8, 4, one, 0), (8, 4, None, 1)):
            st = L.sunny_geneig_launch_chunked(s.handle, 1, 1, 1, 1, 1, total, chunk, streams, ns)
            assert st == sg.STATUS_INVALID_ARG and "launch_chunked" in sg.last_error()
        assert L.sunny_geneig_launch_chunked(s.handle, 0, 0, 0, 0, 0, 0, 4, one, 1) == sg.STATUS_OK


def _phase(cache_dir, handles=1):
    env = dict(os.environ, SUNNY_GENEIG_CACHE_DIR=str(cache_dir))
    out = subprocess.run([sys.executable, os.path.join(REPO, "tests", "cache_phase.py"), str(handles)],
                         env=env, cwd=REPO, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    return json.loads(out.stdout.strip().splitlines()[-1]), out.stderr


def test_cache_layers_and_corruption(tmp_path):
    """Cold start compiles and writes the code object; a new process loads it
    from disk; a second handle in the same process shares the in-process
    module; all give bit-identical W and U. Overwriting the file's ELF magic
    with "XXXX" makes the next process recompile and rewrite it, again with
    the same bits."""
    cold, _ = _phase(tmp_path)
    assert cold["layers"] == ["compile"], cold
    files = sorted(f for f in os.listdir(tmp_path) if f.endswith(".hsaco") and "_fp32_" not in f)
    assert len(files) == 1, files
    path = os.path.join(tmp_path, files[0])
    warm, _ = _phase(tmp_path, handles=2)
    assert warm["layers"] == ["disk", "in-process"], warm
    for r in (cold, warm):
        assert r["info_nonzero"] == 0 and r["eig_rel"] < T.TOL_EIG and r["phase"] < T.TOL_PHASE, r
    assert (warm["w_hash"], warm["u_hash"]) == (cold["w_hash"], cold["u_hash"])
    size = os.path.getsize(path)
    with open(path, "r+b") as f:
        f.write(b"XXXX")
    bad, err = _phase(tmp_path)
    assert bad["layers"] == ["compile"], bad
    assert "invalid code object" in err
    with open(path, "rb") as f:
        assert f.read(4) == b"\x7fELF"
    assert os.path.getsize(path) == size
    assert (bad["w_hash"], bad["u_hash"]) == (cold["w_hash"], cold["u_hash"])
