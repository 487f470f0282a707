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
    # one low-precision solve, and the contraction of the refinement: a
    # wrong capacitance row or level term shows up here long before the
    # refined result misses its target (measured: f16 3.0e-4, 1.1e-7,
    # 4.6e-11; bf16 2.4e-3, 7.0e-6, 2.3e-8, 8.3e-11)
    first, by = (1e-3, 2) if dtype == "f16" else (1e-2, 3)
    assert errs[0] < first
    assert errs[by] < 1e-9
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


@pytest.mark.parametrize("m,nlvl,rank,n,dtype", [
    (512, 3, 32, 5, "f16"),      # whole matrix in the bottom kernel
    (1024, 4, 20, 3, "f16"),     # one upper level, writes x directly
    (8192, 7, 32, 70, "f16"),    # several column groups in both kernels
    (8192, 6, 32, 16, "bf16"),   # leaf 128
    (8192, 7, 64, 9, "f16"),     # rank 64: narrower bottom kernel
    (4096, 6, 80, 4, "f16"),     # rank above 64: the per-level kernels
    (32768, 9, 32, 8, "f16"),    # 64 and 32 tiles per top segment: 4 and 2 finishing phases
    (65536, 10, 32, 3, "f16"),   # 128 tiles per top segment: 8 finishing phases
])
def test_refined_solve_shapes(cuda, m, nlvl, rank, n, dtype):
    import torch
    print(f"mixed shapes m={m} L={nlvl} r={rank} n={n} {dtype}", flush=True)
    A = PlrTree.random(m, nlvl, rank, seed=43)
    P = D.Plan(A)
    assert P.factor() == 0
    assert P.factor_mixed(dtype) == 0
    B = torch.tensor(np.random.default_rng(6).standard_normal((n, m)), device=cuda)
    X64 = torch.zeros_like(B)
    P.solve("N", B, X64)
    ref = X64.cpu().numpy()
    X = torch.zeros_like(B)
    P.solve_refined(B, X, 1)
    e1 = rel(X.cpu().numpy(), ref)
    X = torch.zeros_like(B)
    P.solve_refined(B, X, 10)
    e10 = rel(X.cpu().numpy(), ref)
    print(f"one solve {e1:.1e}, ten {e10:.1e}", flush=True)
    assert e1 < (2e-3 if dtype == "f16" else 1e-2)
    assert e10 <= 1e-10
    # every column on its own reaches the target too
    for c in range(n):
        assert rel(X[c].cpu().numpy(), ref[c]) <= 1e-10
    P.close()


@pytest.mark.parametrize("pad", [6, 3])
def test_refined_solve_strided_blocks(cuda, pad):
    """B and X inside larger arrays (leading dimension m + pad): an even
    leading dimension reads B in place, an odd one goes through a copy."""
    import torch
    m, n = 4096, 20
    print(f"mixed strided pad={pad}", flush=True)
    A = PlrTree.random(m, 6, 32, seed=44)
    P = D.Plan(A)
    assert P.factor() == 0
    assert P.factor_mixed("f16") == 0
    big = torch.tensor(np.random.default_rng(8).standard_normal((n, m + pad)), device=cuda)
    B = big[:, :m]
    Xbig = torch.full((n, m + pad), 7.0, dtype=torch.float64, device=cuda)
    X = Xbig[:, :m]
    P.solve_refined(B, X, 8)
    X64 = torch.zeros((n, m), dtype=torch.float64, device=cuda)
    P.solve("N", B.contiguous(), X64)
    torch.cuda.synchronize()
    assert rel(X.cpu().numpy(), X64.cpu().numpy()) <= 1e-10
    assert torch.all(Xbig[:, m:] == 7.0)     # padding untouched
    assert torch.equal(big[:, :m], B)        # B is not written
    P.close()


@pytest.mark.parametrize("pad,off", [(6, 1), (4, 1), (5, 1)])
def test_refined_solve_unaligned_x(cuda, pad, off):
    """X starting one element into a row (8 bytes past a 16-byte boundary):
    the fused residual product needs an aligned X and declines, and the
    refinement takes the copy path for the residuals instead of failing."""
    import torch
    m, n = 4096, 9
    print(f"mixed unaligned X pad={pad} off={off}", flush=True)
    A = PlrTree.random(m, 6, 32, seed=45)
    P = D.Plan(A)
    assert P.factor() == 0
    assert P.factor_mixed("f16") == 0
    B = torch.tensor(np.random.default_rng(9).standard_normal((n, m)), device=cuda)
    X64 = torch.zeros_like(B)
    P.solve("N", B, X64)
    for mode in ("iters", "tol"):
        Xbig = torch.full((n, m + pad), 3.0, dtype=torch.float64, device=cuda)
        X = Xbig[:, off:off + m]
        if mode == "iters":
            P.solve_refined(B, X, 6)
        else:
            its, est = P.solve_refined_tol(B, X, tol=1e-10, maxit=8)
            assert est <= 1e-10 and its <= 4
        torch.cuda.synchronize()
        assert rel(X.cpu().numpy(), X64.cpu().numpy()) <= 1e-10, mode
        assert torch.all(Xbig[:, :off] == 3.0) and torch.all(Xbig[:, off + m:] == 3.0)
    P.close()


@pytest.mark.parametrize("dtype,most", [("f16", 3), ("bf16", 4)])
def test_refined_tol_reports_its_error(cuda, orc, dtype, most):
    """The refined solve stops on its own error estimate: the estimate it
    reports is <= tol, the true error (against the FP64 solve, checked
    against the oracle in test_gpu_solve.py) is too, and the estimate is
    not far below the true error."""
    import torch
    m, nlvl, rank, n = 8192, 7, 32, 16
    print(f"refined tol {dtype}", flush=True)
    A = PlrTree.random(m, nlvl, rank, seed=41)
    P = D.Plan(A)
    assert P.factor() == 0
    assert P.factor_mixed(dtype) == 0
    B = np.random.default_rng(5).standard_normal((m, n))
    Bd = torch.tensor(np.ascontiguousarray(B.T), device=cuda)
    X64 = torch.zeros_like(Bd)
    P.solve("N", Bd, X64)
    ref = X64.cpu().numpy()
    for tol, cap in ((1e-10, most), (1e-6, most - 1)):
        X = torch.zeros_like(Bd)
        its, est = P.solve_refined_tol(Bd, X, tol=tol, maxit=10)
        err = rel(X.cpu().numpy(), ref)
        print(f"  tol {tol:.0e}: {its} solves, estimate {est:.2e}, error {err:.2e}", flush=True)
        assert 2 <= its <= cap
        assert est <= tol and err <= tol
        assert err / 2 <= est <= 10 * err
    # an unreachable tolerance: maxit solves, an estimate above it
    X = torch.zeros_like(Bd)
    its, est = P.solve_refined_tol(Bd, X, tol=0.0, maxit=3)
    assert its == 3 and est > 0.0
    # residual through the oracle's FP64 product
    X = torch.zeros_like(Bd)
    P.solve_refined_tol(Bd, X, tol=1e-10, maxit=10)
    r = D.gepdmm("N", "N", A, X.cpu().numpy().T.copy(), n, api=orc) - B
    assert np.linalg.norm(r) / np.linalg.norm(B) < 1e-10
    P.close()


@pytest.mark.parametrize("m,nlvl,rank,ld", [(4096, 6, 80, 4096), (4096, 6, 32, 4097),
                                            (512, 3, 32, 512)])
def test_refined_tol_other_forms(cuda, m, nlvl, rank, ld):
    """The estimate from the other code paths: ranks above 64 (per-level
    kernels and the separate update), an odd leading dimension (B copied),
    and a matrix that fits the bottom kernel alone; a zero column stays
    zero with a zero estimate."""
    import torch
    n = 5
    print(f"refined tol forms m={m} r={rank} ld={ld}", flush=True)
    A = PlrTree.random(m, nlvl, rank, seed=46)
    P = D.Plan(A)
    assert P.factor() == 0
    assert P.factor_mixed("f16") == 0
    Bbig = torch.tensor(np.random.default_rng(10).standard_normal((n, ld)), device=cuda)
    Bbig[2] = 0.0
    B = Bbig[:, :m]
    X64 = torch.zeros((n, m), dtype=torch.float64, device=cuda)
    P.solve("N", B.contiguous(), X64)
    Xbig = torch.zeros((n, ld), dtype=torch.float64, device=cuda)
    X = Xbig[:, :m]
    its, est = P.solve_refined_tol(B, X, tol=1e-10, maxit=10)
    torch.cuda.synchronize()
    err = rel(X.cpu().numpy(), X64.cpu().numpy())
    print(f"  {its} solves, estimate {est:.2e}, error {err:.2e}", flush=True)
    assert est <= 1e-10 and err <= 1e-10
    assert torch.all(X[2] == 0.0)
    P.close()


@pytest.mark.parametrize("m,nlvl,rank,n", [
    (8192, 7, 32, 16),     # bottom kernel + upper levels
    (512, 3, 32, 5),       # the bottom kernel alone
    (8192, 6, 32, 9),      # leaf 128
    (4096, 6, 20, 3),      # rank padded to 32
    (65536, 10, 32, 4),    # the bench shape
])
def test_f16_pairs_refine_in_two_solves(cuda, orc, m, nlvl, rank, n):
    """f16 pair storage (hi + lo planes, three MFMAs per product): one solve
    is good to about 1e-6, so two reach 1e-10; against the FP64 solve and
    the oracle's residual."""
    import torch
    print(f"f16x2 m={m} L={nlvl} r={rank} n={n}", flush=True)
    A = PlrTree.random(m, nlvl, rank, seed=47)
    P = D.Plan(A)
    assert P.factor() == 0
    assert P.factor_mixed("f16x2") == 0
    B = np.random.default_rng(11).standard_normal((m, n))
    Bd = torch.tensor(np.ascontiguousarray(B.T), device=cuda)
    X64 = torch.zeros_like(Bd)
    P.solve("N", Bd, X64)
    ref = X64.cpu().numpy()
    X1 = torch.zeros_like(Bd)
    P.solve_refined(Bd, X1, 1)
    e1 = rel(X1.cpu().numpy(), ref)
    X = torch.zeros_like(Bd)
    its, est = P.solve_refined_tol(Bd, X, tol=1e-10, maxit=6)
    err = rel(X.cpu().numpy(), ref)
    print(f"  one solve {e1:.2e}; {its} solves, estimate {est:.2e}, error {err:.2e}", flush=True)
    assert e1 < 2e-5
    assert its == 2 and est <= 1e-10 and err <= 1e-10
    if m <= 8192:
        r = D.gepdmm("N", "N", A, X.cpu().numpy().T.copy(), n, api=orc) - B
        assert np.linalg.norm(r) / np.linalg.norm(B) < 1e-10
    P.close()


@pytest.mark.parametrize("m,nlvl,rank,n,off", [
    (8192, 7, 32, 64, 0),    # the step's shape class: the apply rides on the first residual
    (4096, 6, 32, 16, 0),    # fewer columns than a group
    (4096, 6, 32, 100, 0),   # a chunk above 64 columns: apply made apart
    (4096, 6, 32, 130, 0),   # chunks of 128 and 2: apart, then riding on the residual
    (4096, 6, 32, 64, 1),    # Xa one element off its 16-byte alignment: apply made apart
])
def test_apply_solve_refined(cuda, orc, m, nlvl, rank, n, off):
    """dplrx_apply_solve_refined_tol against its two halves: Ya against the
    FP64 product (and the oracle), X, the solve count and the estimate
    against dplrx_solve_refined_tol on the same B (the same kernels in the
    same order for the solve)."""
    import torch
    print(f"apply+solve m={m} L={nlvl} r={rank} n={n} off={off}", flush=True)
    A = PlrTree.random(m, nlvl, rank, seed=44)
    P = D.Plan(A)
    assert P.factor() == 0
    assert P.factor_mixed("f16x2") == 0
    rng = np.random.default_rng(9)
    Xh, Bh = rng.standard_normal((m, n)), rng.standard_normal((m, n))
    big = torch.zeros(n * m + off, dtype=torch.float64, device=cuda)
    Xa = big[off:].view(n, m)
    Xa.copy_(torch.tensor(np.ascontiguousarray(Xh.T), device=cuda))
    B = torch.tensor(np.ascontiguousarray(Bh.T), device=cuda)
    Ya, X = torch.zeros_like(B), torch.zeros_like(B)
    its, est = P.apply_solve_refined_tol(Xa, Ya, B, X, tol=1e-10, maxit=8)
    Yref, Xref = torch.zeros_like(B), torch.zeros_like(B)
    P.gepdmm("N", 1.0, Xa.contiguous(), 0.0, Yref)
    its2, est2 = P.solve_refined_tol(B, Xref, tol=1e-10, maxit=8)
    torch.cuda.synchronize()
    assert rel(Ya.cpu().numpy(), Yref.cpu().numpy()) < 1e-14
    yo = D.gepdmm("N", "N", A, Xh, n, api=orc)
    assert rel(Ya.cpu().numpy().T, yo) < 1e-12
    # (the correction norms behind the estimate are summed by atomics: the
    # estimates agree to rounding, not bits)
    assert its == its2 and abs(est - est2) <= 1e-9 * est2 and est <= 1e-10
    assert rel(X.cpu().numpy(), Xref.cpu().numpy()) <= 1e-14
    X64 = torch.zeros_like(B)
    P.solve("N", B, X64)
    torch.cuda.synchronize()
    assert rel(X.cpu().numpy(), X64.cpu().numpy()) <= 1e-10
    P.close()
    A.free()
