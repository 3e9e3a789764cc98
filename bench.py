"""Benchmark: batched generalized eigensolve H u = lambda S u on one MI355X
through libsunny_geneig.so's C ABI.

Workload (one "step" = one solve of the whole batch): N = 54 complex, batch
65536, inputs H = Q diag(lambda) Q^H (lambda ~ U[-10, 10], 6 pairs split by
1e-8) and S = B B^H + 0.05 N I, generated on the device (synthetic) and
resident in HBM before timing starts.

The headline is the mixed-precision mode (FP64 Cholesky, FP32 reduction /
tridiagonalisation / QL / eigenvectors, FP64 refinement on the matrix cores).
The plain FP64 fused and split modes are timed on the same batch beside it,
and the accuracy of every mode is checked against the 50 N eps bar (residual
and S-orthogonality, 1-norms). The CPU baseline is LAPACK zhegv over a
bounded sample of the same matrices on the host cores, which also gives the
mixed mode's eigenvalue error against zhegv.

A serving-loop measurement follows (continuous batching through the
library's session; Poisson arrivals at 70% of the measured peak; p50/p99).

Prints one JSON line. Usage: python bench.py [--gpus 1] [--steps K]
[--warmup W] [--batch B] [--n N] [--no-cpu] [--no-fp64] [--no-serve]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0
EPS = 2.0 ** -52


def flops_per_solve(n: int) -> float:
    """LAWN 41 count of the zhegv path: potrf 4/3, hegst 4, hetrd 16/3, Q 16/3,
    QL with vectors 12, back-trsm 4 (x N^3) = 32 N^3 real flops."""
    return 32.0 * n ** 3


def refine_flops_per_solve(n: int, iters: int) -> float:
    """Algorithmic complex multiply-adds of the refinement kernel, 8 real
    flops each: per step X = M V, Y = V^H (X - V diag(lambda)) and V E
    (3 N^3; the Gram entries of close pairs are dot products, not counted),
    then U = L^-H V by back substitution with the Cholesky factor (N^3 / 2,
    the count of a triangular product). Before the last step Y and V E run
    on the FP32 matrix cores. (M = W H W^H, W = L^-1, is formed by its own
    kernel beforehand.)"""
    return 8.0 * (0.5 + 3 * iters) * n ** 3


def bytes_per_solve(n: int) -> float:
    """Compulsory HBM bytes: read H, S; write U, W, info."""
    return 48.0 * n * n + 8.0 * n + 4.0


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(H, S, n, sample: int, threads: int):
    """LAPACK zhegv (the reference's live oracle call) over `sample` matrices:
    1 warm-up + 5 timed runs, median; plus a 1-core figure."""
    from oracle import oracle as O
    provider = O.load_lapack()
    Hs = H[:sample].cpu().numpy()
    Ss = S[:sample].cpu().numpy()
    O.zhegv_batch(Hs[:64], Ss[:64], threads=threads)
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        W, _, info = O.zhegv_batch(Hs, Ss, threads=threads)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    one = min(256, sample)
    t0 = time.perf_counter()
    O.zhegv_batch(Hs[:one], Ss[:one], threads=1)
    t1 = time.perf_counter() - t0
    return {
        "value": sample / med, "unit": "solves/s", "cores": threads, "kind": "port",
        "sample": f"LAPACK zhegv(itype=1,'V','L') via {os.path.basename(provider.split(' ')[0])}, "
                  f"OpenMP over {sample} of the benchmark's N={n} matrices, 1 BLAS thread each, "
                  f"median of 5 runs after 1 warm-up",
        "single_core_solves_per_s": one / t1,
        "info_nonzero": int((info != 0).sum()),
    }, W


def O_zhegv(Hs, Ss, threads: int):
    """zhegv eigenpairs of a host sample (W, U, info)."""
    from oracle import oracle as O
    return O.zhegv_batch(Hs, Ss, threads=threads)


def cpu_baseline_sample(Hs, Ss, threads: int):
    """zhegv eigenvalues of a small host sample (the staged orders' accuracy
    check against the oracle)."""
    from oracle import oracle as O
    W, _, info = O.zhegv_batch(Hs, Ss, threads=threads)
    return info, W


def serving_bench(kernel: str, pools, peak_requests: int, poisson_requests: int, load: float,
                  seed: int, slots: int = 4, policy: str = "fifo", slice_: int = 0, host: bool = False,
                  hbm_budget: int = 1 << 30):
    """Continuous batching through the library's serving session.

    Requests are q-point batches: K pencils, K log-uniform in [16, 4096], one
    order N per request drawn from `pools` (device-resident inputs of the
    benchmark generator; a request uses the first K pencils of its order's
    pool and gets its own output buffers). First a burst of `peak_requests`
    gives the peak sustained rate; then `poisson_requests` arrive as a Poisson
    process at `load` x that rate. Latency = submission to results ready, as
    measured by the session. policy / slice_: the session's admission order
    and slice size. host: inputs and outputs in pinned host memory, staged
    through the session's arena of hbm_budget bytes (admission by footprint
    then limits how many requests are in flight)."""
    import numpy as np
    import torch
    import sunny_gen_eigensolver_amd as sg
    rng = np.random.default_rng(seed)
    orders = sorted(pools)

    def draw(count):
        ks = np.exp(rng.uniform(np.log(16), np.log(4096), size=count)).astype(int)
        ns = rng.choice(orders, size=count)
        return list(zip(ns.tolist(), ks.tolist()))

    ss = sg.ServingSession(orders, kernel=kernel, max_count=4096, slots=slots, hbm_budget=hbm_budget,
                           policy=policy, slice=slice_)
    if host:
        pools = {n: (H.cpu().pin_memory(), S.cpu().pin_memory()) for n, (H, S) in pools.items()}
        outs = {n: (torch.empty((4096, n), dtype=torch.float64).pin_memory(),
                    torch.empty((4096, n, n), dtype=torch.complex128).pin_memory(),
                    torch.zeros((4096,), dtype=torch.int32).pin_memory()) for n in orders}

    def submit(n, k):
        H, S = pools[n]
        if host:
            # Requests share their order's pinned output area: only latency is
            # measured here, and every request's results are written in full.
            W, U, info = (t[:k] for t in outs[n])
            return ss.submit(H[:k], S[:k], W, U, info)
        W = torch.empty((k, n), dtype=torch.float64, device=H.device)
        U = torch.empty((k, n, n), dtype=torch.complex128, device=H.device)
        info = torch.zeros((k,), dtype=torch.int32, device=H.device)
        return ss.submit(H[:k], S[:k], W, U, info)

    # Warm-up, then each order alone on the idle GPU at a few sizes (the
    # service time no schedule can beat), then the peak rate from a burst.
    for n in orders:
        ss.wait(submit(n, 64))
    solo_k = (16, 64, 256, 1024, 4096)
    solo = {n: [min(ss.wait(submit(n, k))[0] for _ in range(2)) for k in solo_k] for n in orders}
    burst = draw(peak_requests)
    t0 = time.perf_counter()
    ids = [submit(n, k) for n, k in burst]
    for i in ids:
        ss.wait(i)
    peak_s = time.perf_counter() - t0
    peak_rate = sum(k for _, k in burst) / peak_s                 # solves/s
    # Poisson arrivals at load x peak.
    reqs = draw(poisson_requests)
    mean_k = float(np.mean([k for _, k in reqs]))
    lam = load * peak_rate / mean_k                                # requests/s
    gaps = rng.exponential(1.0 / lam, size=len(reqs))
    ids = []
    t0 = time.perf_counter()
    t_next = t0
    for (n, k), g in zip(reqs, gaps):
        t_next += g
        while time.perf_counter() < t_next:
            pass
        ids.append(submit(n, k))
    lats = [ss.wait(i)[0] for i in ids]
    span = time.perf_counter() - t0
    st = ss.stats()
    ss.close()
    lats = np.array(lats)
    # Each request's latency is at least its own service time on an idle GPU
    # (interpolated in log K from the solo runs): the p99 of those times is a
    # lower bound on the p99 latency of any schedule for this request mix.
    iso = np.array([np.interp(np.log(k), np.log(solo_k), solo[n]) for n, k in reqs])
    return {
        "mode": kernel, "orders": orders, "requests": len(reqs), "slots": slots, "policy": policy,
        "slice": slice_, "buffers": "pinned host, staged through the arena" if host else "device",
        "hbm_budget_bytes": hbm_budget if host else None, "arena_peak_bytes": st["hbm_peak_bytes"],
        "peak_solves_per_s": peak_rate, "offered_load": load,
        "offered_solves_per_s": lam * mean_k,
        "achieved_solves_per_s": sum(k for _, k in reqs) / span,
        "p50_ms": float(np.percentile(lats, 50)), "p99_ms": float(np.percentile(lats, 99)),
        "max_ms": float(lats.max()), "max_concurrent": st["max_concurrent"],
        "p99_lower_bound_ms": float(np.percentile(iso, 99)), "p50_lower_bound_ms": float(np.percentile(iso, 50)),
        "solo_ms": {str(n): dict(zip(map(str, solo_k), solo[n])) for n in orders},
        "note": "K log-uniform in [16, 4096] per request, N uniform over the served orders; " +
                ("pinned host inputs and outputs (PCIe copies inside the latency); " if host else
                 "device-resident inputs and outputs (no PCIe in the latency); ") +
                f"{slots} slots; latency from submission to results ready",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--n", type=int, default=54)
    ap.add_argument("--seed", type=lambda s: int(s, 0), default=0x5EED0002)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 fused / split modes")
    ap.add_argument("--cpu-sample", type=int, default=4096)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop measurement")
    ap.add_argument("--uneven-tuning", choices=("suggested", "autotune"), default="suggested",
                    help="workgroup size per uneven-batch bucket: fixed rule or the timed sweep")
    ap.add_argument("--no-uneven", action="store_true",
                    help="skip the uneven batch (200000 pencils, N log-uniform over 12..128)")
    ap.add_argument("--uneven-count", type=int, default=200000)
    ap.add_argument("--no-sunny", action="store_true",
                    help="skip the H = Sigma (diagonal) measurement")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the staged-kernel orders (N=96 batch 16384, N=192 batch 2048)")
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-slots", type=int, default=4, help="streams (concurrent requests) of the session")
    ap.add_argument("--serve-only-lds", action="store_true",
                    help="serve only the LDS-resident orders (24, 54) of the request mix")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="launch one device copy of this many MiB first (HBM counter calibration)")
    a = ap.parse_args()

    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: every matrix is independent, so the batched solve "
              "shards by contiguous batch ranges with no exchange; run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import sunny_gen_eigensolver_amd as sg
    from sunny_gen_eigensolver_amd import inputs

    dev = torch.device("cuda", 0)
    n, B = a.n, a.batch
    t0 = time.perf_counter()
    H, S = inputs.spectrum_batch(n, B, a.seed, device=dev)
    torch.cuda.synchronize()
    log(f"generated {B} x N={n} inputs on device in {time.perf_counter() - t0:.1f}s")

    if a.copy_probe_mib > 0:
        # An elementwise kernel (not a DMA copy) that reads and writes exactly
        # copy_probe_mib MiB each: the reference point for HBM byte counters.
        src = torch.ones(a.copy_probe_mib * (1 << 20) // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    W = torch.empty((B, n), dtype=torch.float64, device=dev)
    U = torch.empty((B, n, n), dtype=torch.complex128, device=dev)
    info = torch.zeros((B,), dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()

    def timed(solver, steps, warmup):
        for _ in range(warmup):
            sg.run_kernel(solver, H, S, W, U, info, stream=stream)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            sg.run_kernel(solver, H, S, W, U, info, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps   # ms per step

    def stages(solver):
        """Per-kernel durations of one more launch, from HIP events the library
        records on the launch stream around each kernel."""
        solver.set_stage_timing(True)
        sg.run_kernel(solver, H, S, W, U, info, stream=stream)
        st = solver.stage_times()
        solver.set_stage_timing(False)
        return st

    def accuracy():
        res, orth = inputs.quality(H, S, W, U)
        return {"max_residual_over_N_eps": res.max().item() / (n * EPS),
                "max_s_orth_over_N_eps": orth.max().item() / (n * EPS),
                "bar_over_N_eps": 50.0, "info_nonzero": int((info != 0).sum().item())}

    modes = {}
    W_fp64 = None
    if not a.no_fp64:
        for kernel in ("unified", "hybrid"):
            s = sg.Solver(n, kernel=kernel)
            ms = timed(s, a.steps, a.warmup)
            acc = accuracy()
            st = stages(s)
            if kernel == "unified":
                W_fp64 = W.clone()
            modes[kernel] = {"ms_per_step": ms, "solves_per_s": B / (ms * 1e-3), "accuracy": acc,
                             "stages_ms": {k: v for k, v in st}, "block_dim_x": s.block_dim_x}
            log(f"{kernel}: {ms:.3f} ms/step, {B / (ms * 1e-3):.0f} solves/s, "
                f"resid {acc['max_residual_over_N_eps']:.2f} Ne, s_orth {acc['max_s_orth_over_N_eps']:.2f} Ne")
            s.close()

    mixed = sg.Solver(n, kernel="mixed")
    t_wall = time.perf_counter()
    ms = timed(mixed, a.steps, a.warmup)
    wall = time.perf_counter() - t_wall
    acc = accuracy()
    st = stages(mixed)
    iters = int(os.environ.get("SUNNY_GENEIG_REFINE_ITERS", "2"))
    value = B / (ms * 1e-3)
    log(f"mixed: {ms:.3f} ms/step, {value:.0f} solves/s, resid {acc['max_residual_over_N_eps']:.2f} Ne, "
        f"s_orth {acc['max_s_orth_over_N_eps']:.2f} Ne; stages " +
        ", ".join(f"{k} {v:.2f}" for k, v in st))
    if W_fp64 is not None:
        acc["max_eig_rel_diff_vs_fp64_fused"] = (
            ((W - W_fp64).abs() / W_fp64.abs().clamp(min=1.0)).max().item())

    # Roofline of the path (the algorithmic work of a zhegv solve, 32 N^3
    # real flops, refinement not credited) over the whole step: the mixed
    # step's kernels run back to back as one hipGraph replay, so the step time
    # is the sum of their launch durations (timed on the launch stream).
    path_tflops = B * flops_per_solve(n) / (ms * 1e-3) / 1e12
    traffic = None
    tpath = os.path.join(REPO, "profiles", f"traffic_mixed_n{n}_b{B}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        if tr.get("batch") == B and tr.get("n") == n:
            traffic = tr.get("hbm_bytes_per_step")
    # The refinement kernel's own matrix-core efficiency (its GEMM flops are
    # extra work of the mixed form and are not credited to the path above).
    refine_ms = dict(st).get("geneig_mx_refine")
    refine = None
    if refine_ms:
        rf = refine_flops_per_solve(n, iters)
        refine = {"kernel": "geneig_mx_refine", "ms": refine_ms, "gemm_flops_per_solve": rf,
                  "achieved_TFLOPs": B * rf / (refine_ms * 1e-3) / 1e12,
                  "frac_of_fp64_mfma_peak": B * rf / (refine_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                  "note": "the refinement's complex products (per step M V, V^H R, V E; then "
                          "U = L^-H V by back substitution, N^3/2), 8 real flops per complex multiply-add, "
                          "over the kernel's launch time; before the last step V^H R and V E run on the "
                          "FP32 matrix cores, so the fraction is of the FP64 peak as a reference"}
    dom_name, dom_ms = max(st, key=lambda kv: kv[1])

    out = {
        "metric": "geneig solves/s at N=54 complex, mixed+refined vs FP64 at <=50Ne resid; p99 ms",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64+c128",
        "data": "synthetic",
        "config": {"workload": f"mixed-precision zhegv (FP64 Cholesky, FP32 tridiagonalisation and "
                               f"eigenvectors, {iters} FP64 refinement steps), N={n} complex, batch {B}, "
                               f"H=Q diag(lambda) Q^H (6 pairs split 1e-8), S=BB^H+0.05N I",
                   "n": n, "batch": B, "seed": hex(a.seed), "mode": "mixed", "refine_iters": iters},
        "accuracy": acc,
        "stages_ms": {k: v for k, v in st},
        "roofline": {"bound": "mfma", "achieved": path_tflops, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": path_tflops / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": "mixed step: " + " + ".join(k for k, _ in st),
                     "note": f"32 N^3 = {flops_per_solve(n):.0f} algorithmic flops per solve (the LAPACK "
                             f"zhegv path, LAWN 41 counting; refinement not credited) x batch / the whole "
                             f"step's time; the step's kernels run back to back (one hipGraph replay), "
                             f"timed with HIP events on the launch stream; FP64 peak (vector = matrix); "
                             f"traffic = HBM bytes of all kernels of one step (rocprofv3 FETCH_SIZE + "
                             f"WRITE_SIZE, calibrated)"},
        "dominant_kernel": {"kernel": dom_name, "ms": dom_ms, "share_of_step": dom_ms / ms},
        "refine_gemm": refine,
        "hbm": {"compulsory_GBs": B * bytes_per_solve(n) / (ms * 1e-3) / 1e9,
                "frac_of_peak": B * bytes_per_solve(n) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "fp64_modes": modes,
        "wall_s_timed_region": wall,
    }
    if "unified" in modes:
        out["mixed_vs_fp64_fused"] = value / modes["unified"]["solves_per_s"]
        out["mixed_vs_fp64_split"] = value / modes["hybrid"]["solves_per_s"]
    mixed.close()

    # Sunny's call pattern (H = Sigma = diag(I, -I), the paraunitary metric;
    # S positive definite) on the same S: the mixed mode through
    # sunny_geneig_launch with Sigma written out in full, then through
    # sunny_geneig_launch_hdiag with its diagonal (M formed without H).
    if not a.no_sunny:
        sig = torch.cat([torch.ones(n // 2), -torch.ones(n - n // 2)]).to(torch.float64).to(dev)
        Hs = torch.diag(sig).to(torch.complex128).expand(B, n, n).contiguous()
        ms_d = {}
        with sg.Solver(n, kernel="mixed") as s3:
            for form in ("full", "diagonal"):
                def go():
                    if form == "full":
                        sg.run_kernel(s3, Hs, S, W, U, info, stream=stream)
                    else:
                        sg.run_kernel_hdiag(s3, sig, S, W, U, info, stream=stream)
                for _ in range(a.warmup):
                    go()
                torch.cuda.synchronize()
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for _ in range(a.steps):
                    go()
                e1.record(stream)
                torch.cuda.synchronize()
                ms_d[form] = e0.elapsed_time(e1) / a.steps
                res, orth = inputs.quality(Hs, S, W, U)
                ms_d[form + "_acc"] = {"max_residual_over_N_eps": res.max().item() / (n * EPS),
                                       "max_s_orth_over_N_eps": orth.max().item() / (n * EPS),
                                       "info_nonzero": int((info != 0).sum().item())}
                s3.set_stage_timing(True)
                go()
                ms_d[form + "_stages"] = dict(s3.stage_times())
                s3.set_stage_timing(False)
        out["sunny_sigma"] = {
            "workload": f"mixed mode, H = Sigma = diag(I_{n // 2}, -I_{n - n // 2}), S = the benchmark's S, "
                        f"N={n}, batch {B}",
            "full_h_ms_per_step": ms_d["full"], "diagonal_h_ms_per_step": ms_d["diagonal"],
            "full_h_solves_per_s": B / (ms_d["full"] * 1e-3),
            "diagonal_h_solves_per_s": B / (ms_d["diagonal"] * 1e-3),
            "speedup": ms_d["full"] / ms_d["diagonal"],
            "accuracy": {"full": ms_d["full_acc"], "diagonal": ms_d["diagonal_acc"]},
            "stages_ms": {"full": ms_d["full_stages"], "diagonal": ms_d["diagonal_stages"]}}
        log(f"sunny Sigma/S: full H {ms_d['full']:.2f} ms, diagonal H {ms_d['diagonal']:.2f} ms "
            f"({ms_d['full'] / ms_d['diagonal']:.2f}x); accuracy {ms_d['diagonal_acc']}")
        del Hs
        torch.cuda.empty_cache()

    # Mixed precision at the staged orders (FP64 Cholesky / products /
    # refinement around FP32 tridiagonalisation and eigenvectors) beside the
    # FP64 staged pipeline on the same inputs.
    large, large_samples, uneven_samples = {}, {}, {}
    if not a.no_large:
        for n2, B2 in ((96, 16384), (192, 2048)):
            H2, S2 = inputs.spectrum_batch(n2, B2, a.seed + n2, device=dev)
            W2 = torch.empty((B2, n2), dtype=torch.float64, device=dev)
            U2 = torch.empty((B2, n2, n2), dtype=torch.complex128, device=dev)
            I2 = torch.zeros((B2,), dtype=torch.int32, device=dev)
            rec, W64 = {}, None
            for kernel in ("unified", "mixed"):
                s2 = sg.Solver(n2, kernel=kernel)
                for _ in range(max(1, a.warmup)):
                    sg.run_kernel(s2, H2, S2, W2, U2, I2, stream=stream)
                torch.cuda.synchronize()
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for _ in range(a.steps):
                    sg.run_kernel(s2, H2, S2, W2, U2, I2, stream=stream)
                e1.record(stream)
                torch.cuda.synchronize()
                ms2 = e0.elapsed_time(e1) / a.steps
                res, orth = inputs.quality(H2[:512], S2[:512], W2[:512], U2[:512])
                s2.set_stage_timing(True)
                sg.run_kernel(s2, H2, S2, W2, U2, I2, stream=stream)
                st2 = s2.stage_times()
                s2.set_stage_timing(False)
                s2.close()
                r = {"ms_per_step": ms2, "solves_per_s": B2 / (ms2 * 1e-3),
                     "max_residual_over_N_eps": res.max().item() / (n2 * EPS),
                     "max_s_orth_over_N_eps": orth.max().item() / (n2 * EPS),
                     "info_nonzero": int((I2 != 0).sum().item()),
                     "frac_of_fp64_peak": B2 * flops_per_solve(n2) / (ms2 * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                     "stages_ms": {k: v for k, v in st2}}
                if kernel == "unified":
                    W64 = W2.clone()
                else:
                    r["max_eig_rel_diff_vs_fp64_mode"] = (
                        ((W2 - W64).abs() / W64.abs().clamp(min=1.0)).max().item())
                large_samples[(n2, kernel)] = W2[:32].cpu().numpy()
                rec["fp64" if kernel == "unified" else "mixed"] = r
                log(f"N={n2} B={B2} {kernel}: {ms2:.2f} ms/step, {B2 / (ms2 * 1e-3):.0f} solves/s, resid "
                    f"{r['max_residual_over_N_eps']:.2f} Ne, s_orth {r['max_s_orth_over_N_eps']:.2f} Ne")
            rec["mixed_vs_fp64"] = rec["mixed"]["solves_per_s"] / rec["fp64"]["solves_per_s"]
            rec["batch"] = B2
            large_samples[(n2, "inputs")] = (H2[:32].cpu().numpy(), S2[:32].cpu().numpy())
            large[str(n2)] = rec
            del H2, S2, W2, U2, I2, W64
            torch.cuda.empty_cache()
        out["large_orders"] = large

    # Uneven batch: pencils of orders drawn log-uniform over the integers
    # 12..128 with cond(S) log-uniform in [1e1, 1e4], packed back to back in
    # draw order, solved by one call of the uneven-batch path (LDS-resident
    # buckets up to 64, staged buckets above).
    if not a.no_uneven:
        import ctypes
        import numpy as np
        rng = np.random.default_rng(a.seed)
        cnt = a.uneven_count
        orders = np.floor(np.exp(rng.uniform(np.log(12), np.log(129), size=cnt))).astype(np.int64)
        orders = np.clip(orders, 12, 128)
        conds = np.exp(rng.uniform(np.log(1e1), np.log(1e4), size=cnt))
        nsq = orders * orders
        moff = np.concatenate([[0], np.cumsum(nsq)[:-1]])
        tot = int(nsq.sum())
        t0 = time.perf_counter()
        Hv = torch.empty((tot,), dtype=torch.complex128, device=dev)
        Sv = torch.empty((tot,), dtype=torch.complex128, device=dev)
        for n3 in np.unique(orders):
            idx = np.nonzero(orders == n3)[0]
            Hn, Sn = inputs.spectrum_batch(int(n3), len(idx), a.seed + 1000 + int(n3), device=dev,
                                           cond_s=torch.tensor(conds[idx], device=dev))
            dst = (torch.tensor(moff[idx], device=dev)[:, None]
                   + torch.arange(int(n3) * int(n3), device=dev)[None, :]).reshape(-1)
            Hv[dst] = Hn.reshape(-1)
            Sv[dst] = Sn.reshape(-1)
            del Hn, Sn, dst
        Wv = torch.empty((int(orders.sum()),), dtype=torch.float64, device=dev)
        Uv = torch.empty((tot,), dtype=torch.complex128, device=dev)
        Iv = torch.zeros((cnt,), dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        log(f"uneven: generated {cnt} pencils (sum N^2 = {tot}) in {time.perf_counter() - t0:.1f}s")
        t_vs = time.perf_counter()
        vs = sg.VarlenSolver(128, tuning=a.uneven_tuning)
        log(f"uneven: {a.uneven_tuning} buckets ready in {time.perf_counter() - t_vs:.1f}s, "
            f"matrices per workgroup {vs.bucket_waves}")
        arr = (ctypes.c_int * cnt)(*orders.tolist())
        L = sg.lib()
        sh = sg._stream_handle(stream)

        def vlaunch():
            st_ = L.sunny_geneig_varlen_launch(vs.handle, Hv.data_ptr(), Sv.data_ptr(), Wv.data_ptr(),
                                               Uv.data_ptr(), Iv.data_ptr(), arr, cnt, sh)
            if st_ != sg.STATUS_OK:
                raise RuntimeError(sg.last_error())
        for _ in range(max(1, a.warmup)):
            vlaunch()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(a.steps):
            vlaunch()
        e1.record(stream)
        torch.cuda.synchronize()
        msv = e0.elapsed_time(e1) / a.steps
        # Residual and S-orthogonality per order on a sample of each order.
        woff = np.concatenate([[0], np.cumsum(orders)[:-1]])
        worst_r, worst_o = 0.0, 0.0
        uneven_samples = {}
        for n3 in (12, 33, 54, 64, 65, 96, 128):
            idx = np.nonzero(orders == n3)[0][:64]
            if len(idx) == 0:
                continue
            gi = (torch.tensor(moff[idx], device=dev)[:, None]
                  + torch.arange(n3 * n3, device=dev)[None, :])
            wi = (torch.tensor(woff[idx], device=dev)[:, None] + torch.arange(n3, device=dev)[None, :])
            res, orth = inputs.quality(Hv[gi].view(-1, n3, n3), Sv[gi].view(-1, n3, n3), Wv[wi],
                                       Uv[gi].view(-1, n3, n3))
            worst_r = max(worst_r, res.max().item() / (n3 * EPS))
            worst_o = max(worst_o, orth.max().item() / (n3 * EPS))
            uneven_samples[n3] = (Hv[gi].view(-1, n3, n3).cpu().numpy(), Sv[gi].view(-1, n3, n3).cpu().numpy(),
                                  res.max().item() / (n3 * EPS), orth.max().item() / (n3 * EPS))
        flops = float(sum(flops_per_solve(int(x)) for x in orders))
        out["uneven"] = {
            "count": cnt, "orders": "log-uniform integers 12..128", "cond_s": "log-uniform [1e1, 1e4]",
            "ms_per_step": msv, "solves_per_s": cnt / (msv * 1e-3),
            "frac_of_fp64_peak": flops / (msv * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
            "compulsory_GBs": float((48 * nsq + 8 * orders + 4).sum()) / (msv * 1e-3) / 1e9,
            "share_above_64": float((orders > 64).mean()),
            "info_nonzero": int((Iv != 0).sum().item()),
            "max_residual_over_N_eps_sample": worst_r, "max_s_orth_over_N_eps_sample": worst_o,
            "buckets": vs.buckets, "tuning": a.uneven_tuning, "bucket_waves": vs.bucket_waves,
            "note": "one call per step: host bucket sort and descriptor upload included; sample = up "
                    "to 64 pencils of each of the orders 12, 33, 54, 64, 65, 96, 128",
        }
        log(f"uneven: {msv:.1f} ms/step, {cnt / (msv * 1e-3):.0f} solves/s, resid {worst_r:.2f} Ne, "
            f"s_orth {worst_o:.2f} Ne, info!=0 {out['uneven']['info_nonzero']}")
        vs.close()
        del Hv, Sv, Wv, Uv, Iv
        torch.cuda.empty_cache()

    if not a.no_serve:
        # Serving: one order per request from {24, 54, 96, 150} (the last two
        # on the staged kernels), inputs from the same generator, resident in
        # HBM; then the LDS-resident orders alone.
        pools = {54: (H[:4096], S[:4096])}
        for n3 in (24, 96, 150):
            pools[n3] = inputs.spectrum_batch(n3, 4096, a.seed + n3, device=dev)
        serving = {}
        # FP64 form: the split mode, the fastest FP64 path at these orders.
        runs = [("mixed", (24, 54, 96, 150)), ("hybrid", (24, 54, 96, 150)), ("mixed", (24, 54)),
                ("hybrid", (24, 54))]
        if a.serve_only_lds:
            runs = runs[2:]
        for kernel, orders in runs:
            r = serving_bench(kernel, {o: pools[o] for o in orders}, 200, a.serve_requests, 0.7, a.seed,
                              a.serve_slots)
            form = "fp64_split" if kernel == "hybrid" else kernel
            key = form if len(orders) == 4 else f"{form}_orders_24_54"
            serving[key] = r
            log(f"serving {key}: peak {r['peak_solves_per_s']:.0f} solves/s; at 70%: "
                f"p50 {r['p50_ms']:.2f} ms p99 {r['p99_ms']:.2f} ms, achieved {r['achieved_solves_per_s']:.0f}; "
                f"idle-GPU service p99 {r['p99_lower_bound_ms']:.2f} ms")
        out["serving"] = serving
        out["p99_ms"] = serving["mixed" if "mixed" in serving else "mixed_orders_24_54"]["p99_ms"]
        del pools

    if not a.no_cpu:
        # The host cores this GPU's share of the box allows (16 per GPU; the
        # machine's own count is reported beside it).
        try:
            visible = len(os.sched_getaffinity(0))
        except Exception:
            visible = os.cpu_count() or 1
        threads = min(int(os.environ.get("OMP_NUM_THREADS", "16") or 16), visible)
        # The mixed result for the sample, recomputed (W holds the last mode's).
        m2 = sg.Solver(n, kernel="mixed")
        sg.run_kernel(m2, H, S, W, U, info, stream=stream)
        torch.cuda.synchronize()
        m2.close()
        cb, Wl = cpu_baseline(H, S, n, min(a.cpu_sample, B), threads)
        import numpy as np
        Wg = W[:Wl.shape[0]].cpu().numpy()
        out["accuracy"]["max_eig_rel_err_vs_zhegv_sample"] = float(
            (np.abs(Wg - Wl) / np.maximum(np.abs(Wl), 1.0)).max())
        cb["cpu_model"] = cpu_model()
        cb["host_cpus_visible"] = visible
        out["cpu_baseline"] = cb
        # The uneven batch's sample: zhegv's own residual and S-orthogonality
        # on the same pencils (cond(S) up to 1e4 lifts both above 50 N eps).
        if uneven_samples and "uneven" in out:
            zr, zo, per = 0.0, 0.0, {}
            for n3, (Hs, Ss, r_ours, o_ours) in sorted(uneven_samples.items()):
                Wz, Uz, _ = O_zhegv(Hs, Ss, threads)
                rz, oz = inputs.quality(torch.from_numpy(Hs).to(dev), torch.from_numpy(Ss).to(dev),
                                        torch.from_numpy(Wz).to(dev), torch.from_numpy(Uz).to(dev))
                rz, oz = rz.max().item() / (n3 * EPS), oz.max().item() / (n3 * EPS)
                zr, zo = max(zr, rz), max(zo, oz)
                per[str(n3)] = {"ours_resid": r_ours, "ours_s_orth": o_ours, "zhegv_resid": rz,
                                "zhegv_s_orth": oz}
            out["uneven"]["zhegv_max_residual_over_N_eps_sample"] = zr
            out["uneven"]["zhegv_max_s_orth_over_N_eps_sample"] = zo
            out["uneven"]["per_order_over_N_eps"] = per
        # The staged orders' samples against zhegv as well.
        for key, rec in large.items():
            n2 = int(key)
            Hs, Ss = large_samples[(n2, "inputs")]
            _, Wl2 = cpu_baseline_sample(Hs, Ss, threads)
            for kernel, tag in (("unified", "fp64"), ("mixed", "mixed")):
                Wg2 = large_samples[(n2, kernel)]
                rec[tag]["max_eig_rel_err_vs_zhegv_sample"] = float(
                    (np.abs(Wg2 - Wl2) / np.maximum(np.abs(Wl2), 1.0)).max())
        log(f"cpu: {cb['value']:.0f} solves/s on {threads} threads, "
            f"{cb['single_core_solves_per_s']:.0f} on 1")

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
