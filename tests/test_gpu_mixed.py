"""Mixed-precision (f16 / bf16 factors, FP32 accumulation on MFMA) solves with
FP64-residual iterative refinement: error against the FP64 factored solve
(itself checked against the oracle in test_gpu_solve.py) must fall below the
1e-10 relative target."""
import numpy as np
import pytest

import dplr_amd as D
from dplr_amd.tree import PlrTree

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("dtype,iters", [("f16", 6), ("bf16", 9)])
def test_refined_solve_reaches_1e10(cuda, orc, dtype, iters):
    import torch
    m, nlvl, rank, n = 8192, 7, 32, 16
    print(f"mixed {dtype} m={m}", flush=True)
    A = PlrTree.random(m, nlvl, rank, seed=41)
    B = np.random.default_rng(5).standard_normal((m, n))
    P = D.Plan(A)
    assert P.factor() == 0
    assert P.factor_mixed(dtype) == 0
    Bd = torch.tensor(np.ascontiguousarray(B.T), device=cuda)
    X64 = torch.zeros_like(Bd)
    P.solve("N", Bd, X64)
    errs = []
    for k in range(1, iters + 1):
        Xd = torch.zeros_like(Bd)
        P.solve_refined(Bd, Xd, k)
        torch.cuda.synchronize()
        errs.append(rel(Xd.cpu().numpy(), X64.cpu().numpy()))
    print(dtype, ["%.1e" % e for e in errs], flush=True)
    assert errs[0] < 5e-2                      # one low-precision solve
    assert errs[-1] <= 1e-10                   # refined to the target
    # residual through the oracle's FP64 matvec
    X = Xd.cpu().numpy().T
    r = D.gepdmm("N", "N", A, X, n, api=orc) - B
    assert np.linalg.norm(r) / np.linalg.norm(B) < 1e-10
    P.close()


def test_refined_residual_report(cuda):
    import torch
    m = 4096
    A = PlrTree.random(m, 6, 32, seed=42)
    P = D.Plan(A)
    assert P.factor_mixed("f16") == 0
    B = torch.randn(3, m, dtype=torch.float64, device=cuda)
    X = torch.zeros_like(B)
    res = P.solve_refined(B, X, 6, with_residual=True)
    assert res.shape == (3,) and np.all(res < 1e-9 * B.abs().max().item())
    P.close()
