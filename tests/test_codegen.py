"""Code-generation guard for the gfx950 kernels (CPU only: hipcc emits the
assembly, nothing runs).

A kernel that solves several matrices per workgroup in a loop (the fused
kernel does) once had, at some orders, per-lane LDS addresses kept live
across that loop and saved into accumulation registers (v_accvgpr_write) in a
basic block whose exec mask was still narrowed by the `if (lane == 0)` region
that had just ended; the matching restore then ran with every lane active, so
the second matrix of a workgroup was factored from wrong addresses (zhegv INFO
n + 2 on an HPD S). This test checks every order the package ships warm for
the pattern: a save that is still unrestored when `s_or_b64 exec, exec, ...`
widens the mask."""
import os
import re
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(REPO, "sunny_gen_eigensolver_amd", "csrc", "kernels", "geneig_kernels.hip")
KERNELS = ("geneig_fused", "geneig_mx_front", "geneig_split_ql", "geneig_ql_full", "geneig_mx_back",
           "geneig_mx_backtri")
ORDERS = sorted(set(range(4, 65, 4)) | {27, 54})


def _block_width(n):
    return 64 if n <= 16 else 128 if n <= 32 else 256


def _kernels(asm):
    out, cur = {}, None
    for line in asm.split("\n"):
        m = re.match(r"^(geneig_\w+):", line)
        if m:
            cur = m.group(1)
            out[cur] = []
        elif cur:
            out[cur].append(line)
            if "s_endpgm" in line:
                cur = None
    return out


def narrowed_saves(lines):
    """Blocks holding an unrestored VGPR->AGPR save when exec is widened."""
    hits, block, saves = [], None, []
    for line in lines:
        if re.match(r"^(\.LBB\d+_\d+:|; %bb\.\d+:)", line):
            block, saves = line.split()[0 if line.startswith(".") else 1], []
            continue
        t = line.strip()
        m = re.match(r"v_accvgpr_write_b32 (a\d+), (v\d+)", t)
        if m:
            saves.append((m.group(1), m.group(2)))
        m = re.match(r"v_accvgpr_read_b32 (v\d+), (a\d+)", t)
        if m:
            saves = [x for x in saves if x != (m.group(2), m.group(1))]
        if re.match(r"s_or_b64 exec, exec, s\[\d+:\d+\]", t) and saves:
            hits.append((block, list(saves)))
            saves = []
        elif t.startswith(("s_branch", "s_cbranch")):
            saves = []
    return hits


def _compile(n, tmp):
    out = os.path.join(tmp, f"k{n}.s")
    subprocess.run(["hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                    "-Xclang", "-packed-fp32-ops", f"-DGENEIG_N={n}", f"-DGENEIG_NT={_block_width(n)}",
                    "-S", "--cuda-device-only", SRC, "-o", out], check=True, capture_output=True)
    return n, open(out).read()


def test_detector_flags_the_pattern():
    bad = ["; %bb.1:", "\tv_accvgpr_write_b32 a5, v126", "\ts_or_b64 exec, exec, s[60:61]"]
    ok = ["; %bb.1:", "\tv_accvgpr_write_b32 a5, v126", "\tv_accvgpr_read_b32 v126, a5",
          "\ts_or_b64 exec, exec, s[60:61]"]
    assert narrowed_saves(bad) and not narrowed_saves(ok)


@pytest.mark.skipif(subprocess.run(["which", "hipcc"], capture_output=True).returncode != 0,
                    reason="hipcc not available")
def test_no_saves_under_narrowed_exec(tmp_path):
    with ThreadPoolExecutor(max(1, min(8, len(os.sched_getaffinity(0))))) as ex:
        results = list(ex.map(lambda n: _compile(n, str(tmp_path)), ORDERS))
    bad = []
    for n, asm in results:
        ks = _kernels(asm)
        assert set(KERNELS) <= set(ks), f"N={n}: kernels missing from assembly"
        for k in KERNELS:
            for block, saves in narrowed_saves(ks[k]):
                bad.append(f"N={n} {k} {block}: {saves}")
    assert not bad, "\n".join(bad)
