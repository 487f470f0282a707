"""Orders beyond the benchmark's: N = 2^20 (14 levels, rank 16) and 2^21
(15 levels, rank 8), where 32-bit index arithmetic over N x (sum of ranks)
or N x columns would be the first thing to break.  The products against
the oracle's apply (src/dplrpdmm.c:454, src/dplrpmv.c:262), the FP64 solve
by its residual through the oracle's apply (the oracle's explicit inverse
is out of reach at these orders), the refined mixed solve against the FP64
one and its own estimate."""
import numpy as np
import pytest

import dplr_amd as D

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("m,nlvl,rank,n", [(1 << 20, 14, 16, 4), (1 << 21, 15, 8, 2)])
def test_large_orders(cuda, orc, m, nlvl, rank, n):
    import torch
    print(f"large order m={m} L={nlvl} r={rank} n={n}", flush=True)
    A = D.PlrTree.random(m, nlvl, rank, 4.0, 7)
    P = D.Plan(A)
    rng = np.random.default_rng(3)
    Xh, Bh = rng.standard_normal((m, n)), rng.standard_normal((m, n))
    X = torch.tensor(np.ascontiguousarray(Xh.T), device=cuda)
    Y = torch.zeros_like(X)
    P.gepdmm("N", 1.0, X, 0.0, Y)
    y1 = torch.zeros(m, dtype=torch.float64, device=cuda)
    P.gepmv("N", 1.0, X[0].contiguous(), 0.0, y1)
    torch.cuda.synchronize()
    Yo = D.gepdmm("N", "N", A, Xh, n, api=orc)
    assert rel(Y.cpu().numpy().T, Yo) < 1e-14
    assert rel(y1.cpu().numpy(), Yo[:, 0]) < 1e-14
    assert P.factor() == 0
    B = torch.tensor(np.ascontiguousarray(Bh.T), device=cuda)
    Z = torch.zeros_like(B)
    P.solve("N", B, Z)
    torch.cuda.synchronize()
    R = D.gepdmm("N", "N", A, Z.cpu().numpy().T.copy(), n, api=orc) - Bh
    assert np.linalg.norm(R) / np.linalg.norm(Bh) < 1e-13
    assert P.factor_mixed("f16x2") == 0
    Z2 = torch.zeros_like(B)
    its, est = P.solve_refined_tol(B, Z2, tol=1e-10, maxit=8)
    torch.cuda.synchronize()
    assert est <= 1e-10 and rel(Z2.cpu().numpy(), Z.cpu().numpy()) <= 1e-10
    P.close()
    A.free()
