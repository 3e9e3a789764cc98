"""GPU parity of the uneven-batch path (sunny_geneig_varlen_*): pencils of
different orders in one call, each checked against FP64 LAPACK zhegv on the
same matrix with the fixed-order bars (eig_rel and phase < 1e-10, residual
and S-orthogonality <= 50 N eps), plus zhegv's INFO codes for a non-HPD S in
the mix, and the alone-versus-mixed property: a matrix's result does not
depend on what else is in the call (bit-identical: every matrix is solved by
its own workgroups of the bucket's kernels whatever its neighbours, and the
calls here stay below the batch size at which the QL switches to its
rotation log)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as T

pytestmark = pytest.mark.gpu


def _gen_mix(orders, seed):
    """Per-matrix (H, S) from the reference generator, one seed per entry."""
    Hs, Ss = [], []
    for k, n in enumerate(orders):
        H, S = O.gen_reference(n, 1, seed + k, "random")
        Hs.append(H[0])
        Ss.append(S[0])
    return Hs, Ss


def _solve(vs, Hs, Ss, orders):
    import torch
    import sunny_gen_eigensolver_amd as sg
    dev = torch.device("cuda", 0)
    dH = sg.pack([torch.from_numpy(np.ascontiguousarray(h)) for h in Hs]).to(dev)
    dS = sg.pack([torch.from_numpy(np.ascontiguousarray(s)) for s in Ss]).to(dev)
    W, U, info = vs.run(dH, dS, orders)
    torch.cuda.synchronize()
    Wl = [w.cpu().numpy() for w in sg.unpack_vectors(W, orders)]
    Ul = [u.cpu().numpy() for u in sg.unpack_matrices(U, orders)]
    return Wl, Ul, info.cpu().numpy()


def _check_groups(Hs, Ss, W, U, info, orders, tag):
    by_n = {}
    for k, n in enumerate(orders):
        by_n.setdefault(n, []).append(k)
    for n, ks in sorted(by_n.items()):
        H = np.stack([Hs[k] for k in ks])
        S = np.stack([Ss[k] for k in ks])
        Wg = np.stack([W[k] for k in ks])
        Ug = np.stack([U[k] for k in ks])
        ig = info[ks]
        Wl, Ul, il = O.zhegv_batch(H, S, threads=8)
        assert (il == 0).all()
        T.assert_parity(H, S, Wg, Ug, ig, Wl, Ul, tag=f"{tag} n={n} x{len(ks)}")


@pytest.fixture(scope="module")
def vs64():
    import sunny_gen_eigensolver_amd as sg
    s = sg.VarlenSolver(64)
    yield s
    s.close()


def test_default_buckets(vs64):
    assert vs64.buckets == list(range(4, 65, 4))


def test_mixed_orders_vs_lapack(vs64):
    rng = np.random.default_rng(7)
    orders = [1, 2, 3, 4, 5, 12, 16, 17, 31, 32, 33, 54, 63, 64]
    orders += [int(x) for x in rng.integers(1, 65, size=50)]
    rng.shuffle(orders)
    Hs, Ss = _gen_mix(orders, 5000)
    W, U, info = _solve(vs64, Hs, Ss, orders)
    _check_groups(Hs, Ss, W, U, info, orders, "varlen")


def test_log_uniform_mix_vs_lapack(vs64):
    """The uneven-batch workload's order law (log-uniform), restricted to the
    LDS-resident range 12..64."""
    rng = np.random.default_rng(11)
    orders = [int(x) for x in np.exp(rng.uniform(np.log(12), np.log(64.999), size=300))]
    Hs, Ss = _gen_mix(orders, 9000)
    W, U, info = _solve(vs64, Hs, Ss, orders)
    _check_groups(Hs, Ss, W, U, info, orders, "varlen log-uniform")


def test_alone_equals_mixed(vs64):
    orders = [54, 7, 33, 54, 16, 64, 21, 54, 9]
    Hs, Ss = _gen_mix(orders, 123)
    W, U, info = _solve(vs64, Hs, Ss, orders)
    for k in (0, 2, 5, 6):
        Wa, Ua, ia = _solve(vs64, [Hs[k]], [Ss[k]], [orders[k]])
        assert ia[0] == info[k] == 0
        assert np.array_equal(Wa[0], W[k]), f"matrix {k}: eigenvalues differ alone vs mixed"
        assert np.array_equal(Ua[0], U[k]), f"matrix {k}: eigenvectors differ alone vs mixed"


def test_fixed_order_solver_agrees(vs64):
    """A bucket-order matrix solved by the uneven path matches the fused
    fixed-order kernel to the parity bars."""
    meta, H, S = T.fixture_inputs("n32_b8_s42_random")
    orders = [32] * 8
    W, U, info = _solve(vs64, list(H), list(S), orders)
    Wf, Uf, _ = T.run_gpu(H, S)
    T.assert_parity(H, S, np.stack(W), np.stack(U), info, Wf, Uf, tag="varlen vs fused n=32")


def test_non_hpd_in_mix(vs64):
    """zhegv INFO = n + k for the first non-positive leading minor k of S; the
    other matrices in the call are unaffected."""
    orders = [12, 20, 12, 40]
    Hs, Ss = _gen_mix(orders, 77)
    bad = np.eye(20, dtype=np.complex128)
    bad[6, 6] = -1.0
    Ss[1] = bad.T.copy()
    W, U, info = _solve(vs64, Hs, Ss, orders)
    _, _, il = O.zhegv_batch(Hs[1][None], Ss[1][None])
    assert int(info[1]) == 20 + 7 == int(il[0])
    assert np.isnan(W[1]).all() and (U[1] == 0).all()
    good = [0, 2, 3]
    _check_groups([Hs[k] for k in good], [Ss[k] for k in good], [W[k] for k in good],
                  [U[k] for k in good], info[good], [orders[k] for k in good], "varlen beside non-HPD")


def test_multiple_rounds_per_bucket():
    """More matrices than one round of a bucket holds (65536 at order 4): the
    run is split over several gather / solve / scatter rounds sharing one
    padded staging area."""
    import sunny_gen_eigensolver_amd as sg
    n = 4
    count = 65536 + 37
    H, S = O.gen_reference(n, count, 31337, "random")
    orders = [3 if k % 5 == 0 else 4 for k in range(count)]
    Hs = [H[k][:o, :o].copy() if o < n else H[k] for k, o in enumerate(orders)]
    Ss = [S[k][:o, :o].copy() if o < n else S[k] for k, o in enumerate(orders)]
    with sg.VarlenSolver(8, buckets=[4, 8]) as vs:
        W, U, info = _solve(vs, Hs, Ss, orders)
    _check_groups(Hs, Ss, W, U, info, orders, "varlen multi-round")


def test_rejects_order_above_max():
    import torch
    import sunny_gen_eigensolver_amd as sg
    with sg.VarlenSolver(16) as vs:
        dev = torch.device("cuda", 0)
        x = torch.zeros(17 * 17, dtype=torch.complex128, device=dev)
        with pytest.raises(sg.GeneigError) as ei:
            vs.run(x, x, [17])
        assert ei.value.status == sg.STATUS_INVALID_ARG


@pytest.mark.parametrize("bucket", list(range(4, 65, 4)))
def test_every_default_bucket_every_slot(vs64, bucket):
    """Each default bucket with the workgroups of its QL grids full (one
    matrix per wave): every wave slot holds a matrix of the bucket's own order
    or just below it (padded), so the second and later matrices of a
    workgroup are covered at every order."""
    import sunny_gen_eigensolver_amd as sg
    nw = (64 if bucket <= 16 else 128 if bucket <= 32 else 256) // 64
    lo = max(1, bucket - 3)
    orders = [bucket if k % 2 == 0 else max(lo, bucket - 1 - (k // 2) % 3) for k in range(4 * nw)]
    Hs, Ss = _gen_mix(orders, 40000 + 100 * bucket)
    W, U, info = _solve(vs64, Hs, Ss, orders)
    _check_groups(Hs, Ss, W, U, info, orders, f"bucket {bucket}")


@pytest.mark.parametrize("n", list(range(4, 65, 4)))
def test_fused_at_bucket_orders(n):
    """The fixed-order fused, split and mixed modes at every bucket order (their
    code objects ship warm with the buckets), several matrices per workgroup;
    for the mixed mode's refinement these orders cover one to four 16-blocks
    and partial blocks."""
    H, S = O.gen_reference(n, 9, 50000 + n, "random")
    Wl, Ul, il = O.zhegv_batch(H, S, threads=8)
    assert (il == 0).all()
    for kernel in ("unified", "hybrid", "mixed"):
        W, U, info = T.run_gpu(H, S, kernel=kernel)
        T.assert_parity(H, S, W, U, info, Wl, Ul, tag=f"{kernel} n={n}")


# ---------------------------------------------------------------------------
# Buckets above 64 (the staged kernels on a padded copy of the bucket).
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def vs128():
    import sunny_gen_eigensolver_amd as sg
    s = sg.VarlenSolver(128)
    yield s
    s.close()


def test_default_buckets_128(vs128):
    assert vs128.buckets == list(range(4, 65, 4)) + list(range(72, 129, 8))


def test_log_uniform_12_128_vs_lapack(vs128):
    """The uneven-batch workload's order law: log-uniform over 12..128, so
    about 30% of the pencils fall in the staged buckets."""
    rng = np.random.default_rng(13)
    orders = [int(x) for x in np.exp(rng.uniform(np.log(12), np.log(128.999), size=120))]
    orders += [65, 72, 73, 100, 127, 128]
    rng.shuffle(orders)
    Hs, Ss = _gen_mix(orders, 21000)
    W, U, info = _solve(vs128, Hs, Ss, orders)
    _check_groups(Hs, Ss, W, U, info, orders, "varlen 12..128")


def test_large_bucket_alone_equals_mixed(vs128):
    orders = [100, 7, 128, 54, 65, 100, 90]
    Hs, Ss = _gen_mix(orders, 31000)
    W, U, info = _solve(vs128, Hs, Ss, orders)
    for k in (0, 2, 4, 6):
        Wa, Ua, ia = _solve(vs128, [Hs[k]], [Ss[k]], [orders[k]])
        assert ia[0] == info[k] == 0
        assert np.array_equal(Wa[0], W[k]), f"matrix {k}: eigenvalues differ alone vs mixed"
        assert np.array_equal(Ua[0], U[k]), f"matrix {k}: eigenvectors differ alone vs mixed"


def test_large_bucket_non_hpd(vs128):
    """INFO = n + k in a staged bucket (the padded order's code mapped back to
    the pencil's own order), W NaN and U zero; neighbours unaffected."""
    orders = [90, 90, 100, 77]
    Hs, Ss = _gen_mix(orders, 41000)
    Ss = [s.copy() for s in Ss]
    bad = {1: 0, 2: 70}
    for k, p in bad.items():
        Sm = np.eye(orders[k], dtype=np.complex128)
        Sm[p, p] = -1.0
        Ss[k] = Sm.T.copy()
    W, U, info = _solve(vs128, Hs, Ss, orders)
    for k, p in bad.items():
        assert int(info[k]) == orders[k] + p + 1, (k, info[k])
        assert np.isnan(W[k]).all() and (U[k] == 0).all()
    good = [0, 3]
    _check_groups([Hs[k] for k in good], [Ss[k] for k in good], [W[k] for k in good], [U[k] for k in good],
                  info[good], [orders[k] for k in good], "varlen beside non-HPD")


def test_autotuned_buckets_vs_lapack():
    """tuning="autotune": every bucket takes what the split-mode solver's
    timed sweep picks for its order (the same choice as
    Solver(n, tuning="autotune", kernel="hybrid"), read back from the
    persisted table), and the results stay within the parity bars. A bucket
    whose sweep picked the HBM-staged kernels (orders from 48) reports one
    matrix per workgroup."""
    import sunny_gen_eigensolver_amd as sg
    buckets = [16, 24, 32, 48]
    with sg.VarlenSolver(48, buckets=buckets, tuning="autotune") as vs:
        assert vs.buckets == buckets
        for o, waves in zip(buckets, vs.bucket_waves):
            s = sg.Solver(o, tuning="autotune", kernel="hybrid")
            want = 1 if s.staged else s.block_dim_x // 64
            assert waves == want, (o, waves, s.block_dim_x, s.staged)
            s.close()
        rng = np.random.default_rng(23)
        orders = [int(x) for x in rng.integers(9, 49, size=40)]
        Hs, Ss = _gen_mix(orders, 8100)
        W, U, info = _solve(vs, Hs, Ss, orders)
    _check_groups(Hs, Ss, W, U, info, orders, "varlen autotuned")


def _gen_c3(orders, conds, seed):
    """The uneven workload's law: per pencil the benchmark generator with
    S = Q_S diag(sigma) Q_S^H, sigma geometric from 1 to cond(S)."""
    Hs, Ss = [], []
    for k, (n, c) in enumerate(zip(orders, conds)):
        H, S = O.gen_spectrum(int(n), 1, seed + k, cond_s=float(c))
        Hs.append(H[0])
        Ss.append(S[0])
    return Hs, Ss


def test_c3_law_vs_zhegv(vs128):
    """Orders log-uniform over 12..128 (LDS-resident and staged buckets in one
    call) with cond(S) log-uniform in [1e1, 1e4]: each pencil against zhegv
    on the same pencil, with the conditioning-aware bars (residual and
    S-orthogonality <= max(50 N eps, 2 x zhegv's own), eigenvalues within
    1e-12 ||H||_2)."""
    rng = np.random.default_rng(0xC3)
    cnt = 240
    orders = np.clip(np.floor(np.exp(rng.uniform(np.log(12), np.log(129), size=cnt))), 12, 128).astype(int)
    orders[:4] = [12, 64, 65, 128]                   # both sides of the LDS / staged boundary
    conds = np.exp(rng.uniform(np.log(1e1), np.log(1e4), size=cnt))
    conds[:4] = 1e4
    Hs, Ss = _gen_c3(orders.tolist(), conds, 0xC3000)
    W, U, info = _solve(vs128, Hs, Ss, orders.tolist())
    by_n = {}
    for k, n in enumerate(orders.tolist()):
        by_n.setdefault(n, []).append(k)
    # Groups of neighbouring orders keep zhegv batches reasonably sized.
    for n, ks in sorted(by_n.items()):
        H = np.stack([Hs[k] for k in ks])
        S = np.stack([Ss[k] for k in ks])
        T.assert_parity_conditioned(H, S, np.stack([W[k] for k in ks]), np.stack([U[k] for k in ks]),
                                    info[ks], tag=f"C3 n={n} x{len(ks)}")


def test_graph_replay_of_uneven_calls(vs128):
    """A call is captured once per (buffers, orders) into one hipGraph that
    holds every grid, staged buckets included; repeated calls replay it with
    bit-identical results, a call with other orders in between gets its own
    graph and does not disturb the first, and direct launches agree."""
    import torch
    import sunny_gen_eigensolver_amd as sg
    rng = np.random.default_rng(0x6A)
    o1 = [int(x) for x in rng.integers(12, 129, size=40)] + [70, 128, 64]
    o2 = [int(x) for x in rng.integers(1, 129, size=25)]
    H1, S1 = _gen_mix(o1, 7100)
    H2, S2 = _gen_mix(o2, 7200)
    dev = torch.device("cuda", 0)
    packs = {}
    for tag, (Hs, Ss, oo) in (("a", (H1, S1, o1)), ("b", (H2, S2, o2))):
        dH = sg.pack([torch.from_numpy(np.ascontiguousarray(h)) for h in Hs]).to(dev)
        dS = sg.pack([torch.from_numpy(np.ascontiguousarray(s_)) for s_ in Ss]).to(dev)
        W = torch.empty((sum(oo),), dtype=torch.float64, device=dev)
        U = torch.empty((sum(x * x for x in oo),), dtype=torch.complex128, device=dev)
        I = torch.zeros((len(oo),), dtype=torch.int32, device=dev)
        packs[tag] = (dH, dS, W, U, I, oo)
    c0, r0 = vs128.graph_stats()
    outs = []
    for tag in ("a", "a", "b", "a", "b"):
        dH, dS, W, U, I, oo = packs[tag]
        W.fill_(np.nan)
        U.zero_()
        vs128.run(dH, dS, oo, W, U, I)
        torch.cuda.synchronize()
        outs.append((tag, W.cpu().numpy().copy(), U.cpu().numpy().copy(), I.cpu().numpy().copy()))
    c1, r1 = vs128.graph_stats()
    assert (c1 - c0, r1 - r0) == (2, 3), (c1 - c0, r1 - r0)
    ref = {t: o for t, *o in outs[:1] + outs[2:3]}
    for tag, W, U, I in outs:
        assert np.array_equal(W, ref[tag][0]) and np.array_equal(U, ref[tag][1]) and np.array_equal(I, ref[tag][2])
    vs128.set_graphs(False)
    try:
        dH, dS, W, U, I, oo = packs["a"]
        vs128.run(dH, dS, oo, W, U, I)
        torch.cuda.synchronize()
        assert np.array_equal(W.cpu().numpy(), ref["a"][0]) and np.array_equal(U.cpu().numpy(), ref["a"][1])
    finally:
        vs128.set_graphs(True)
    Wl = [w for w in sg.unpack_vectors(torch.from_numpy(ref["a"][0]), o1)]
    Ul = [u.numpy() for u in sg.unpack_matrices(torch.from_numpy(ref["a"][1]), o1)]
    _check_groups(H1, S1, [w.numpy() for w in Wl], Ul, ref["a"][2], o1, "graph replay")
