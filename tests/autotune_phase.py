"""One autotuning phase, run in a fresh process (the in-process table would
otherwise answer): create a solver of order N in MODE with
TUNING_AUTOTUNE, report the configuration it runs (workgroup size, matrices
per workgroup, LDS-resident or HBM-staged blocking), then solve a batch of
the benchmark generator's pencils with it and check them against zhegv.
Prints one JSON line; the sweep's own log goes to stderr.

Usage: python tests/autotune_phase.py N MODE   (cache directory and table
from SUNNY_GENEIG_CACHE_DIR)."""
import json
import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)


def main():
    import sunny_gen_eigensolver_amd as sg
    from oracle import oracle as O
    from tests import helpers as T
    n, mode = int(sys.argv[1]), sys.argv[2]
    with sg.Solver(n, kernel=mode, tuning="autotune") as s:
        cfg = {"block_dim": s.block_dim_x, "batches_per_block": s.batches_per_block, "staged": s.staged}
        H, S = O.gen_spectrum(n, 24, 0x70E + n)
        W, U, info = T.run_gpu(H, S, solver=s)
    Wl, Ul, _ = O.zhegv_batch(H, S, threads=8)
    T.assert_parity(H, S, W, U, info, Wl, Ul, tag=f"autotuned {mode} n={n} {cfg}")
    table = open(os.path.join(sg.cache_dir(), "autotune.tsv")).read().split("\n")
    print(json.dumps({"config": cfg, "table": [t for t in table if t]}), flush=True)


if __name__ == "__main__":
    main()
