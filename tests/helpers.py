"""Shared test helpers: golden fixtures, input generation, GPU runs through
the C ABI, and the reference's comparison metrics (via the oracle)."""
import json
import os

import numpy as np

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
EPS = 2.0 ** -52
TOL_EIG = 1e-10        # reference bar (tests/static/test_full_n54.cpp:144-182)
TOL_PHASE = 1e-10
DEGEN_GAP = 1e-6
RESID_FACTOR = 50.0    # bar: residual and S-orthogonality <= 50 N eps (1-norms)
EIG_NORM_BAR = 1e-12   # eigenvalue bar of the benchmark's pencils: |dlambda| <= 1e-12 ||H||_2 per matrix


def fixtures_index():
    with open(os.path.join(GOLDEN, "fixtures.json")) as f:
        return json.load(f)


def golden_arrays():
    return np.load(os.path.join(GOLDEN, "fixtures.npz"), allow_pickle=False)


def fixture_inputs(name):
    """Regenerate a reference fixture's H, S with the oracle's generator."""
    meta = fixtures_index()[name]
    mode = "random" if meta["s_mode"] == "random" else "near"
    H, S = O.gen_reference(meta["n"], meta["batch"], meta["seed"], mode)
    return meta, H, S


def run_gpu(H, S, kernel="unified", solver=None, n=None):
    """Solve a host batch on cuda:0 through the C ABI; returns W, U, info (numpy)."""
    import torch
    import sunny_gen_eigensolver_amd as sg
    b, n = H.shape[0], H.shape[1]
    own = solver is None
    if own:
        solver = sg.Solver(n, kernel=kernel)
    dev = torch.device("cuda", 0)
    dH = torch.from_numpy(np.ascontiguousarray(H)).to(dev)
    dS = torch.from_numpy(np.ascontiguousarray(S)).to(dev)
    W, U, info = sg.run_kernel(solver, dH, dS)
    torch.cuda.synchronize()
    out = W.cpu().numpy(), U.cpu().numpy(), info.cpu().numpy()
    if own:
        solver.close()
    return out


def metrics(H, S, W, U, W_ref=None, U_ref=None):
    return O.check_batch(H, S, W, U, W_ref, U_ref, DEGEN_GAP, threads=os.cpu_count() or 1)


def norm2(A):
    """Spectral norm of each matrix of a batch (column-major storage)."""
    return np.linalg.norm(np.asarray(A), 2, axis=(1, 2))


def eig_abs_ratio(H, W, W_ref):
    """max over each matrix of |W - W_ref| / (1e-12 ||H||_2): <= 1 passes the
    eigenvalue bar."""
    d = np.abs(np.asarray(W) - np.asarray(W_ref)).max(axis=1)
    return d / (EIG_NORM_BAR * norm2(H))


def assert_parity(H, S, W, U, info, W_ref, U_ref, tag=""):
    """The reference's bars (eig_rel, phase < 1e-10, tests/static/
    test_full_n54.cpp:144-182), residual and S-orthogonality <= 50 N eps, and
    |W - W_ref| <= 1e-12 ||H||_2 per matrix."""
    n = H.shape[1]
    assert (info == 0).all(), f"{tag}: nonzero info {info[info != 0][:8]}"
    m = metrics(H, S, W, U, W_ref, U_ref)
    bar = RESID_FACTOR * n * EPS
    er = eig_abs_ratio(H, W, W_ref)
    print(f"{tag}: max eig_rel {m[:,0].max():.3e} phase {m[:,1].max():.3e} "
          f"|dlam|/(1e-12 |H|_2) {er.max():.3e} "
          f"resid/(N eps) {m[:,3].max()/(n*EPS):.3f} s_orth/(N eps) {m[:,4].max()/(n*EPS):.3f}")
    assert m[:, 0].max() < TOL_EIG, f"{tag}: eig_rel {m[:,0].max():.3e}"
    assert m[:, 1].max() < TOL_PHASE, f"{tag}: phase {m[:,1].max():.3e}"
    assert er.max() <= 1.0, f"{tag}: |dlambda| {er.max():.3e} x 1e-12 ||H||_2"
    assert m[:, 3].max() <= bar, f"{tag}: residual {m[:,3].max():.3e} > {bar:.3e}"
    assert m[:, 4].max() <= bar, f"{tag}: S-orthogonality {m[:,4].max():.3e} > {bar:.3e}"
    return m


def assert_parity_conditioned(H, S, W, U, info, tag=""):
    """Parity for pencils whose S is ill-conditioned, where LAPACK itself
    misses 50 N eps S-orthogonality: zhegv solves the same pencils, and each
    pencil must meet residual and S-orthogonality <= max(50 N eps, 2 x
    zhegv's own on that pencil), eigenvalues within 1e-12 ||H||_2 of zhegv's,
    and the reference's eig_rel / phase bars. Prints ||H||_2 ||S^-1||_2 eps,
    the scale of the eigenvalue error any backward-stable solver makes."""
    n = H.shape[1]
    assert (info == 0).all(), f"{tag}: nonzero info {info[info != 0][:8]}"
    Wl, Ul, il = O.zhegv_batch(H, S, threads=os.cpu_count() or 1)
    assert (il == 0).all()
    m = metrics(H, S, W, U, Wl, Ul)
    mz = metrics(H, S, Wl, Ul)
    base = RESID_FACTOR * n * EPS
    bar_r = np.maximum(base, 2.0 * mz[:, 3])
    bar_o = np.maximum(base, 2.0 * mz[:, 4])
    er = eig_abs_ratio(H, W, Wl)
    sinv = 1.0 / np.array([np.linalg.eigvalsh(s.T)[0] for s in S])
    scale = (norm2(H) * sinv).max() * EPS
    print(f"{tag}: ||H|| ||S^-1|| eps <= {scale:.2e}; |dlam|/(1e-12 |H|_2) {er.max():.3e}; "
          f"ours resid {m[:,3].max()/(n*EPS):.2f} s_orth {m[:,4].max()/(n*EPS):.2f}, zhegv resid "
          f"{mz[:,3].max()/(n*EPS):.2f} s_orth {mz[:,4].max()/(n*EPS):.2f} (N eps); worst ratio to "
          f"the per-pencil bar: resid {(m[:,3]/bar_r).max():.3f} s_orth {(m[:,4]/bar_o).max():.3f}")
    assert m[:, 0].max() < TOL_EIG, f"{tag}: eig_rel {m[:,0].max():.3e}"
    assert m[:, 1].max() < TOL_PHASE, f"{tag}: phase {m[:,1].max():.3e}"
    assert er.max() <= 1.0, f"{tag}: |dlambda| {er.max():.3e} x 1e-12 ||H||_2"
    assert (m[:, 3] <= bar_r).all(), f"{tag}: residual above max(50 N eps, 2 x zhegv)"
    assert (m[:, 4] <= bar_o).all(), f"{tag}: S-orthogonality above max(50 N eps, 2 x zhegv)"
    return m, mz
