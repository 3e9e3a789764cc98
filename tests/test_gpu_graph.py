"""hipGraph replay of the multi-kernel launches (split, mixed and staged
modes): the first launch on a set of buffers captures the kernels, later
launches on the same buffers replay the graph. Every replay, with other
device work and synchronisations of the program between replays, must give
the same bits as the kernels launched directly; new buffers or a new batch
size are captured anew."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kernel,n", [("hybrid", 54), ("mixed", 54), ("unified", 96), ("mixed", 96)])
def test_replays_match_direct_launches(kernel, n):
    import torch
    import sunny_gen_eigensolver_amd as sg
    dev = torch.device("cuda", 0)
    H, S = O.gen_spectrum(n, 48, 0x5EED0100 + n)
    dH = torch.from_numpy(np.ascontiguousarray(H)).to(dev)
    dS = torch.from_numpy(np.ascontiguousarray(S)).to(dev)
    with sg.Solver(n, kernel=kernel) as direct:
        direct.set_graphs(False)
        W0, U0, i0 = sg.run_kernel(direct, dH, dS)
        torch.cuda.synchronize()
        W0, U0 = W0.cpu().numpy(), U0.cpu().numpy()
        assert direct.graph_stats() == (0, 0)
    with sg.Solver(n, kernel=kernel) as s:
        W = torch.empty((48, n), dtype=torch.float64, device=dev)
        U = torch.empty((48, n, n), dtype=torch.complex128, device=dev)
        info = torch.zeros((48,), dtype=torch.int32, device=dev)
        for rep in range(4):
            W.fill_(0.0)
            U.zero_()
            sg.run_kernel(s, dH, dS, W, U, info)
            junk = torch.randn(1 << 20, device=dev).sum()      # unrelated device work in between
            torch.cuda.synchronize()
            assert float(junk) == float(junk)
            assert np.array_equal(W.cpu().numpy(), W0), f"replay {rep}: eigenvalues differ"
            assert np.array_equal(U.cpu().numpy(), U0), f"replay {rep}: eigenvectors differ"
            assert (info.cpu().numpy() == 0).all()
        assert s.graph_stats() == (1, 3)
        # A different batch size is a new capture.
        W2, U2, _ = sg.run_kernel(s, dH[:16], dS[:16])
        torch.cuda.synchronize()
        assert s.graph_stats()[0] == 2
        assert np.array_equal(W2.cpu().numpy(), W0[:16])
