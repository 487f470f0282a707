"""Parity at the stated sizes of the workloads the library is measured on.

C1: N = 65536, 10 levels, leaf 64, rank 32, 64 right-hand sides; the FP64
    multi-RHS product in both directions against the oracle's dplrgepdmm
    (the reference algorithm, src/dplrpdmm.c:454), and the FP64 factored
    solve through the oracle's residual (src/dplrinv.c:93 is the inverse it
    replaces; the oracle's explicit inverse is too slow at this size).
C2: the same generator at N = 65536 (rank 32) and N = 131072 (rank 64, 11
    levels), 64 right-hand sides, f16 and bf16 factors refined to an
    estimated 1e-10: the error against the FP64 solve (itself pinned by C1's
    residual test) and the oracle's FP64 residual.
"""
import numpy as np
import pytest

import dplr_amd as D
from dplr_amd.tree import PlrTree

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def c1(cuda):
    print("C1: generating N=65536 L=10 r=32", flush=True)
    A = PlrTree.random(65536, 10, 32, 4.0, seed=1)
    P = D.Plan(A)
    yield A, P
    P.close()
    A.free()


@pytest.mark.parametrize("trans,alpha,beta", [("N", 1.0, 0.0), ("T", 0.5, -1.0)])
def test_c1_product_64rhs(c1, orc, trans, alpha, beta):
    """The fused all-level reduction, its finishing pass and the third-form
    expansion at the bench shape, against the oracle; 1e-12 relative."""
    import torch
    A, P = c1
    m, n = 65536, 64
    print(f"C1 product 64 RHS trans={trans}", flush=True)
    rng = np.random.default_rng(21)
    B = rng.standard_normal((m, n))
    C0 = rng.standard_normal((m, n))
    Bd = torch.tensor(B.T.copy(), device="cuda")
    Cd = torch.tensor(C0.T.copy(), device="cuda")
    P.gepdmm(trans, alpha, Bd, beta, Cd)
    torch.cuda.synchronize()
    ref = D.gepdmm(trans, "N", A, B, n, alpha, beta, C0, api=orc)
    assert rel(Cd.cpu().numpy().T, ref) < 1e-12


def test_c1_solve_64rhs(c1, orc):
    """The FP64 factored solve of 64 right-hand sides (bottom kernel,
    reassociated upper levels): normwise residual through the oracle's
    product on every column."""
    import torch
    A, P = c1
    m, n = 65536, 64
    print("C1 FP64 solve 64 RHS", flush=True)
    assert P.factor() == 0
    B = np.random.default_rng(22).standard_normal((m, n))
    Bd = torch.tensor(B.T.copy(), device="cuda")
    Xd = torch.zeros_like(Bd)
    P.solve("N", Bd, Xd)
    torch.cuda.synchronize()
    X = Xd.cpu().numpy().T
    R = D.gepdmm("N", "N", A, X, n, api=orc) - B
    assert np.linalg.norm(R) / np.linalg.norm(B) < 1e-13
    for j in range(n):
        assert np.linalg.norm(R[:, j]) / np.linalg.norm(B[:, j]) < 1e-13, j


C2_POINTS = [(65536, 10, 32, 64), (131072, 11, 64, 16)]


@pytest.fixture(scope="module", params=C2_POINTS, ids=["N65536_r32", "N131072_r64"])
def c2(request, cuda):
    import torch
    m, nlvl, rank, ncheck = request.param
    print(f"C2: generating N={m} L={nlvl} r={rank}", flush=True)
    A = PlrTree.random(m, nlvl, rank, 4.0, seed=7)
    P = D.Plan(A)
    assert P.factor() == 0
    B = np.random.default_rng(23).standard_normal((m, 64))
    Bd = torch.tensor(B.T.copy(), device="cuda")
    X64 = torch.zeros_like(Bd)
    P.solve("N", Bd, X64)
    torch.cuda.synchronize()
    yield A, P, B, Bd, X64.cpu().numpy(), ncheck
    P.close()
    A.free()
    torch.cuda.empty_cache()


@pytest.mark.parametrize("dtype", ["f16", "bf16"])
def test_c2_refined_64rhs(c2, orc, dtype):
    """Refined to an estimated 1e-10 (normwise over the 64 columns): the
    reported estimate, the error against the FP64 solve (normwise, and no
    column more than 2.5 times the target) and the oracle's FP64 residual on
    the first ncheck columns.  A relative error e gives a relative residual
    of about e * |A||x|/|b|; 3e-10 bounds it here."""
    import torch
    A, P, B, Bd, X64, ncheck = c2
    m = B.shape[0]
    print(f"C2 refined {dtype} N={m}", flush=True)
    assert P.factor_mixed(dtype) == 0
    X = torch.zeros_like(Bd)
    its, est = P.solve_refined_tol(Bd, X, tol=1e-10, maxit=8)
    Xh = X.cpu().numpy()
    err = rel(Xh, X64)
    print(f"  {its} solves, estimate {est:.2e}, error vs FP64 solve {err:.2e}", flush=True)
    assert est <= 1e-10
    assert err <= 1e-10
    assert its <= (3 if dtype == "f16" else 4)
    for c in range(64):
        assert rel(Xh[c], X64[c]) <= 2.5e-10, c
    R = D.gepdmm("N", "N", A, np.ascontiguousarray(Xh[:ncheck].T), ncheck, api=orc) - B[:, :ncheck]
    assert np.linalg.norm(R) / np.linalg.norm(B[:, :ncheck]) < 3e-10
