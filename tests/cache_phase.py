"""One phase of the code-object cache test, run in a fresh process (the
in-process layer is per process): create fused-mode solvers for the
n54_b8_s42_random fixture through the C ABI, report each handle's cache
layer, launch the fixture and print one JSON line with the layers, the
FNV-1a-64 hashes of W and U (the reference's bit-identity check,
tests/unified/test_cache_n54.cpp:139-147) and the accuracy against the
fixture's stored eigenpairs.

Usage: python tests/cache_phase.py HANDLES   (cache directory from
SUNNY_GENEIG_CACHE_DIR; HANDLES solvers are created while the first is
alive, so the second and later ones come from the in-process layer)."""
import ctypes
import json
import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)


def fnv1a64(data: bytes) -> int:
    h = 0xcbf29ce484222325
    a = np.frombuffer(data, dtype=np.uint8)
    # Vectorised in blocks would change the order; the byte loop is exact
    # and the buffers are small (W and U of 8 matrices of order 54).
    for b in a.tolist():
        h ^= b
        h = (h * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def main():
    import torch
    import sunny_gen_eigensolver_amd as sg
    from tests import helpers as T
    handles = int(sys.argv[1]) if len(sys.argv) > 1 else 1
    L = sg.lib()
    name = "n54_b8_s42_random"
    meta, H, S = T.fixture_inputs(name)
    g = T.golden_arrays()
    n, b = H.shape[1], H.shape[0]
    hs, layers = [], []
    for _ in range(handles):
        h = L.sunny_geneig_create(n, 0, 1, sg.TUNING_SUGGESTED, sg.CACHE_AUTO)
        if not h:
            raise SystemExit(f"create failed: {sg.last_error()}")
        hs.append(h)
        layers.append(L.sunny_geneig_cache_layer(h).decode())
    dev = torch.device("cuda", 0)
    dH = torch.from_numpy(np.ascontiguousarray(H)).to(dev)
    dS = torch.from_numpy(np.ascontiguousarray(S)).to(dev)
    W = torch.empty((b, n), dtype=torch.float64, device=dev)
    U = torch.empty((b, n, n), dtype=torch.complex128, device=dev)
    info = torch.zeros((b,), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    st = L.sunny_geneig_launch(hs[-1], dH.data_ptr(), dS.data_ptr(), W.data_ptr(), U.data_ptr(),
                               info.data_ptr(), b, 0)
    if st != 0:
        raise SystemExit(f"launch failed: {sg.last_error()}")
    torch.cuda.synchronize()
    Wn, Un, In = W.cpu().numpy(), U.cpu().numpy(), info.cpu().numpy()
    m = T.metrics(H, S, Wn, Un, g[f"{name}__W"], g[f"{name}__U"])
    for h in hs:
        L.sunny_geneig_destroy(h)
    print(json.dumps({"layers": layers, "w_hash": fnv1a64(Wn.tobytes()), "u_hash": fnv1a64(Un.tobytes()),
                      "info_nonzero": int((In != 0).sum()), "eig_rel": float(m[:, 0].max()),
                      "phase": float(m[:, 1].max()), "cache_dir": sg.cache_dir()}), flush=True)


if __name__ == "__main__":
    main()
