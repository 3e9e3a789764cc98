"""The HBM-staged pipeline at orders up to 64 and the autotuner that weighs
it against the LDS-resident kernels (reference: the BlockDim sweep,
src/unified/unified_solver.cpp:649-925, and its test
tests/unified/test_autotune_n54.cpp: a sweep on first construction, a
cache hit on the second, correct results from the winner).

- SUNNY_GENEIG_STAGED=1 forces the staged pipeline at LDS-resident orders:
  parity with zhegv in every mode.
- Autotuning (every mode here, not only the fused one) times the workgroup
  sizes 64 / 128 / 256 (1 / 2 / 4 matrices per workgroup) and, from order 48,
  the staged pipeline; every candidate is gated on the suggested one's
  eigenvalues and eigenvector phases; the winner is persisted per (order,
  arch, mode) and a second process reads it back without a sweep."""
import glob
import json
import os
import shutil
import subprocess
import sys

import pytest

from oracle import oracle as O
from tests import helpers as T

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("kernel,n", [("unified", 54), ("hybrid", 54), ("mixed", 54), ("mixed", 12),
                                      ("unified", 64)])
def test_staged_pipeline_at_lds_orders(kernel, n, monkeypatch):
    import sunny_gen_eigensolver_amd as sg
    monkeypatch.setenv("SUNNY_GENEIG_STAGED", "1")
    H, S = O.gen_spectrum(n, 32, 0x57A6 + n)
    with sg.Solver(n, kernel=kernel) as s:
        assert s.staged and s.batches_per_block == 1 and s.block_dim_x == 256
        W, U, info = T.run_gpu(H, S, solver=s)
    Wl, Ul, _ = O.zhegv_batch(H, S, threads=8)
    T.assert_parity(H, S, W, U, info, Wl, Ul, tag=f"staged {kernel} n={n}")
    monkeypatch.delenv("SUNNY_GENEIG_STAGED")
    with sg.Solver(n, kernel=kernel) as s:
        assert not s.staged


def _phase(cache_dir, n, mode):
    env = dict(os.environ, SUNNY_GENEIG_CACHE_DIR=str(cache_dir))
    env.pop("SUNNY_GENEIG_STAGED", None)
    out = subprocess.run([sys.executable, os.path.join(REPO, "tests", "autotune_phase.py"), str(n), mode],
                         env=env, cwd=REPO, capture_output=True, text=True, timeout=900)
    assert out.returncode == 0, out.stderr[-3000:]
    return json.loads(out.stdout.strip().splitlines()[-1]), out.stderr


@pytest.mark.parametrize("mode", ["mixed", "unified"])
def test_autotune_sweep_then_table(tmp_path, mode):
    import sunny_gen_eigensolver_amd as sg
    n = 48
    # The LDS-resident code objects of this order come warm from the in-tree
    # cache; the staged ones compile here.
    for f in glob.glob(os.path.join(sg.DEFAULT_CACHE_DIR, f"geneig_n{n}_*.hsaco")):
        shutil.copy(f, tmp_path)
    first, log1 = _phase(tmp_path, n, mode)
    swept = [line for line in log1.splitlines() if "[sunny_geneig autotune]" in line and "median" in line]
    # Every candidate ran: three workgroup sizes, LDS-resident, and the staged pipeline.
    assert sum(" lds " in line for line in swept) == 3, log1[-2000:]
    assert sum(" staged " in line for line in swept) == 1, log1[-2000:]
    cfg = first["config"]
    if cfg["staged"]:
        assert cfg["block_dim"] == 256 and cfg["batches_per_block"] == (1 if mode != "unified" else 1)
    else:
        assert cfg["block_dim"] in (64, 128, 256)
        if mode == "unified":
            assert cfg["batches_per_block"] == cfg["block_dim"] // 64
    mode_id = {"unified": 0, "hybrid": 1, "mixed": 2}[mode]
    rows = [r.split() for r in first["table"]]
    assert [str(n), "gfx950", str(mode_id), str(cfg["block_dim"]), str(int(cfg["staged"]))] in rows, rows
    second, log2 = _phase(tmp_path, n, mode)
    assert second["config"] == cfg
    assert "median" not in log2, "the second process must read the table, not sweep"
