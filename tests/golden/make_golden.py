"""Write the golden vectors under tests/golden/ from the reference's fixtures.

Reads the reference's GEIG v1 data files (tests/data/*.bin in the reference
checkout, zhegvd outputs written by tests/common/lapack_reference.cpp) and
stores, per file:
  - the generator parameters (N, B, seed, S mode),
  - SHA-256 of the H and S bytes, so a test can prove the oracle's generator
    reproduces the reference's inputs bit for bit without shipping them,
  - W (float64 eigenvalues, all files),
  - U (complex128 eigenvectors) for the two smallest-footprint files.
Run in the development container only: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
from oracle import oracle as O  # noqa: E402

SRC = "/root/reference/tests/data"
KEEP_U = {"n16_b8_s42_random", "n54_b8_s42_random"}


def main():
    index = {}
    arrays = {}
    for fn in sorted(os.listdir(SRC)):
        if not fn.endswith(".bin"):
            continue
        name = fn[:-4]
        fx = O.geig_read(os.path.join(SRC, fn))
        mode = "random" if name.endswith("random") else "near-identity"
        index[name] = dict(n=fx["n"], batch=fx["batch"], seed=fx["seed"], s_mode=mode,
                           sha256_H=hashlib.sha256(fx["H"].tobytes()).hexdigest(),
                           sha256_S=hashlib.sha256(fx["S"].tobytes()).hexdigest())
        arrays[f"{name}__W"] = fx["W"]
        if name in KEEP_U:
            arrays[f"{name}__U"] = fx["U"]
    with open(os.path.join(HERE, "fixtures.json"), "w") as f:
        json.dump(index, f, indent=1, sort_keys=True)
    np.savez_compressed(os.path.join(HERE, "fixtures.npz"), **arrays)
    print("wrote", len(index), "fixtures")


if __name__ == "__main__":
    main()
