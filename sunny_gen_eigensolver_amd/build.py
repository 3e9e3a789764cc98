"""Build libsunny_geneig.so in-tree and warm its code-object cache.

Steps:
  1. embed kernels/geneig_kernels.hip as a C string (csrc/_gen/kernel_source.inc)
     for the run-time hipRTC compiler;
  2. compile the host runtime + C ABI with g++ against the HIP runtime and
     hipRTC -> sunny_gen_eigensolver_amd/lib/libsunny_geneig.so;
  3. compile the kernel source ahead of time with hipcc for gfx950 (a build
     check that also prints the kernels' register / LDS budgets);
  4. compile the code objects for the common orders through the library's own
     sunny_geneig_precompile into the in-tree cache directory, so a fresh
     machine loads them from disk instead of running hipRTC.

Run: python -m sunny_gen_eigensolver_amd.build [--no-warm]
"""
from __future__ import annotations

import argparse
import ctypes
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
GEN = os.path.join(CSRC, "_gen")
LIBDIR = os.path.join(PKG, "lib")
BUILD = os.path.join(PKG, "_build")
CACHE = os.path.join(PKG, "_cocache")
KERNEL = os.path.join(CSRC, "kernels", "geneig_kernels.hip")
LIB = os.path.join(LIBDIR, "libsunny_geneig.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIP_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

# Orders whose code objects ship warm at every workgroup size: the reference
# fixtures' LDS-resident sizes and the serving sizes. The uneven-batch buckets
# (multiples of 4 up to 64) ship warm at their own workgroup size.
WARM_ORDERS = [16, 24, 32, 48, 54, 64]
BUCKET_ORDERS = list(range(4, 65, 4))
LARGE_ORDERS = list(range(72, 129, 8)) + [150, 192]   # staged kernels (orders above 64)
# The staged pipeline at LDS-resident orders, which the autotuner weighs from
# order 48 (the uneven-batch buckets and the fixtures' orders there).
STAGED_LDS_ORDERS = [48, 52, 54, 56, 60, 64]


def _run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, **kw)


def embed_kernel_source() -> str:
    os.makedirs(GEN, exist_ok=True)
    src = open(KERNEL).read()
    # hipRTC sees one text: the staged kernels' file goes in place of its #include.
    inc = '#include "geneig_large.hip"\n'
    assert src.count(inc) == 1
    src = src.replace(inc, open(os.path.join(os.path.dirname(KERNEL), "geneig_large.hip")).read())
    delim = "GENEIG_SRC"
    assert (")" + delim + '"') not in src
    out = os.path.join(GEN, "kernel_source.inc")
    text = ("// Generated by sunny_gen_eigensolver_amd/build.py from kernels/geneig_kernels.hip\n"
            "// (with kernels/geneig_large.hip inlined).\n"
            f'static const char kGeneigKernelSource[] = R"{delim}(' + src + f'){delim}";\n')
    old = open(out).read() if os.path.exists(out) else None
    if old != text:
        open(out, "w").write(text)
    return out


def build_library() -> str:
    embed_kernel_source()
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(BUILD, exist_ok=True)
    srcs = [os.path.join(CSRC, f) for f in ("codeobj.cpp", "solver.cpp", "varlen.cpp", "serving.cpp", "c_api.cpp")]
    inc = [f"-I{os.path.join(REPO, 'include')}", f"-I{CSRC}"]
    # The no-allocation launch holds its scratch in a file-scope device array,
    # so its translation unit is compiled by hipcc for gfx950.
    noalloc_o = os.path.join(BUILD, "noalloc.o")
    _run(["hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", *HIP_FLAGS, *inc, "-c",
          os.path.join(CSRC, "noalloc.hip"), "-o", noalloc_o])
    tmp = LIB + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-pthread",
          "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", *inc, *srcs, noalloc_o, "-o", tmp,
          f"-L{ROCM}/lib", "-lamdhip64", "-lhiprtc", f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, LIB)
    return LIB


def aot_check() -> None:
    """Ahead-of-time compile of both module precisions. The FP32 build must not
    promote any arithmetic to FP64 (an unsuffixed literal would)."""
    os.makedirs(BUILD, exist_ok=True)
    _run(["hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17", *HIP_FLAGS, "-c", KERNEL,
          "-o", os.path.join(BUILD, "geneig_kernels.o")])
    _run(["hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17", *HIP_FLAGS, "-DGENEIG_FP32",
          "-Werror=double-promotion", "-c", KERNEL, "-o", os.path.join(BUILD, "geneig_kernels_fp32.o")])
    # The staged kernels (orders above 64), FP32 build.
    _run(["hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17", *HIP_FLAGS, "-DGENEIG_FP32", "-DGENEIG_N=192",
          "-Werror=double-promotion", "-c", KERNEL, "-o", os.path.join(BUILD, "geneig_large_fp32.o")])


def _precompile(job):
    n, bd = job
    lib = ctypes.CDLL(LIB)
    lib.sunny_geneig_precompile.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_char_p]
    lib.sunny_geneig_last_error.restype = ctypes.c_char_p
    rc = lib.sunny_geneig_precompile(n, bd, b"gfx950")
    return n, bd, rc, (lib.sunny_geneig_last_error().decode() if rc else "")


def warm_cache(orders=WARM_ORDERS, buckets=BUCKET_ORDERS, large=LARGE_ORDERS, staged=STAGED_LDS_ORDERS) -> None:
    from concurrent.futures import ProcessPoolExecutor
    os.makedirs(CACHE, exist_ok=True)
    # Entries are keyed by a hash of source + options: drop stale ones.
    for f in os.listdir(CACHE):
        if f.endswith(".hsaco") or f == "autotune.tsv":
            os.remove(os.path.join(CACHE, f))
    os.environ["SUNNY_GENEIG_CACHE_DIR"] = CACHE
    jobs = [(n, bd) for n in orders for bd in (64, 128, 256)]
    jobs += [(n, 0) for n in buckets if n not in orders]
    jobs = [(n, 0) for n in large] + [(n, -1) for n in staged] + jobs   # the longest compiles first
    workers = max(1, min(8, len(os.sched_getaffinity(0))))
    with ProcessPoolExecutor(workers) as ex:
        for n, bd, rc, msg in ex.map(_precompile, jobs):
            if rc != 0:
                raise RuntimeError(f"precompile N={n} wg={bd} failed: {msg}")
    print(f"warmed {len(jobs)} code objects", flush=True)


def build(warm: bool = True) -> None:
    build_library()
    aot_check()
    if warm:
        # A child process: the cache directory is read from the environment.
        _run([sys.executable, "-c",
              "import sunny_gen_eigensolver_amd.build as b; b.warm_cache()"], cwd=REPO)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--no-warm", action="store_true")
    a = ap.parse_args()
    build(warm=not a.no_warm)
