"""Single-precision products and solves: FP32 factors with FP32 arithmetic
on FP32 MFMA, and BF16 factor matrices with FP32 arithmetic; single plans in
both directions, batches of matrices of different order, leaf and rank
(apply and solve in both directions, the BF16 solve refined with FP32
residuals), and the FP32 adjoint.  Every result is compared with the
oracle's FP64 one: its dplrgepdmm (src/dplrpdmm.c:454) for products and for
the residual of solves, its explicit inverse (src/dplrinv.c:93) applied to
the right-hand sides where it is small enough.

Tolerances (relative, Frobenius): FP32 products and solves 2e-5 (unit
roundoff 6e-8 times the reduction lengths and the conditioning of these
diagonally dominant matrices), BF16 1.5e-2 (unit roundoff 3.9e-3 on every
factor entry), the refined BF16 solve 2e-5 (FP32 residuals)."""
import numpy as np
import pytest

import dplr_amd as D
from dplr_amd.tree import PlrTree

pytestmark = pytest.mark.gpu

TOL = {"f32": 2e-5, "bf16": 1.5e-2}
SPECS = [(1024, 4, 8), (2048, 5, 16), (4096, 6, 48), (1024, 3, 20), (2048, 3, 33),
         (8192, 5, 12), (512, 1, 40), (4096, 4, 24)]
NRHS = 16


def orc_op(orc, tree, x, trans="N"):
    """The oracle's op(A) x for an (n, m) tensor, as an (n, m) float64 array."""
    xh = x.double().cpu().numpy().T.copy()
    return D.gepdmm(trans, "N", tree, xh, xh.shape[1], api=orc).T


def orc_rel(a, ref):
    a = a.double().cpu().numpy() if hasattr(a, "cpu") else a
    ref = ref.double().cpu().numpy() if hasattr(ref, "cpu") else ref
    return float(np.linalg.norm(a - ref) / np.linalg.norm(ref))


@pytest.fixture(scope="module")
def one(cuda):
    """N = 16384 (leaf 128, 7 levels, rank 32) with FP32 and BF16 factors."""
    t = PlrTree.random(16384, 7, 32, 4.0, 21)
    P = D.Plan(t)
    assert P.factor_single("f32") == 0
    assert P.factor_single("bf16") == 0
    yield t, P
    P.close()
    t.free()


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_single_apply(one, orc, dtype, trans):
    import torch
    t, P = one
    print("single apply", dtype, trans, flush=True)
    g = torch.Generator(device="cpu").manual_seed(2)
    X = torch.randn(NRHS, t.m, generator=g).to("cuda")
    Y = torch.randn(NRHS, t.m, generator=g).to("cuda")
    Y0 = Y.clone()
    P.gepdmm_single(trans, 0.5, X, -2.0, Y, dtype)
    torch.cuda.synchronize()
    ref = 0.5 * orc_op(orc, t, X, trans) - 2.0 * Y0.double().cpu().numpy()
    assert orc_rel(Y, ref) < TOL[dtype]


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_single_solve(one, orc, dtype, trans):
    import torch
    t, P = one
    print("single solve", dtype, trans, flush=True)
    g = torch.Generator(device="cpu").manual_seed(3)
    X = torch.randn(NRHS, t.m, generator=g).to("cuda")
    Z = torch.zeros_like(X)
    for _ in range(3):  # eager, captured, replayed
        Z.zero_()
        P.solve_single(trans, X, Z, dtype)
        torch.cuda.synchronize()
        assert orc_rel(orc_op(orc, t, Z, trans), X) < TOL[dtype]


@pytest.fixture(scope="module")
def batch(cuda):
    import torch
    trees = [PlrTree.random(m, L, r, 4.0, 300 + k) for k, (m, L, r) in enumerate(SPECS)]
    plans = [D.Plan(t) for t in trees]
    for P in plans:
        assert P.factor_single("f32") == 0
        assert P.factor_single("bf16") == 0
    B = D.Batch(plans)
    g = torch.Generator(device="cpu").manual_seed(6)
    X = [torch.randn(NRHS, m, generator=g).to("cuda") for (m, _, _) in SPECS]
    yield trees, plans, B, X
    B.close()
    for p in plans:
        p.close()
    for t in trees:
        t.free()


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_batch_apply(batch, orc, dtype, trans):
    import torch
    trees, plans, B, X = batch
    print("batch apply", dtype, trans, flush=True)
    Y = [torch.zeros_like(x) for x in X]
    for _ in range(2):
        B.gepdmm(X, Y, prec=dtype, trans=trans)
        torch.cuda.synchronize()
        for k, (T, x, y) in enumerate(zip(trees, X, Y)):
            assert orc_rel(y, orc_op(orc, T, x, trans)) < TOL[dtype], SPECS[k]


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_batch_solve(batch, orc, dtype, trans):
    import torch
    trees, plans, B, X = batch
    print("batch solve", dtype, trans, flush=True)
    inv_ref = {}
    for k, (m, _, _) in enumerate(SPECS):
        if m <= 1024:
            Binv, info = D.geinv(trees[k], 0.0, api=orc)
            assert info == 0
            inv_ref[k] = orc_op(orc, Binv, X[k], trans)
            Binv.free()
    Z = [torch.zeros_like(x) for x in X]
    for rep in range(3):  # eager build, capture, replay
        for z in Z:
            z.zero_()
        B.solve(X, Z, prec=dtype, trans=trans)
        torch.cuda.synchronize()
        for k, (T, x, z) in enumerate(zip(trees, X, Z)):
            if k in inv_ref:
                assert orc_rel(z, inv_ref[k]) < TOL[dtype], (rep, SPECS[k])
            assert orc_rel(orc_op(orc, T, z, trans), x) < TOL[dtype], (rep, SPECS[k])


def test_batch_refined_bf16(batch, orc):
    """BF16 solves refined with FP32 residuals reach FP32 accuracy; every
    solve contracts the error by about the BF16 solve's own error."""
    import torch
    trees, plans, B, X = batch
    print("batch refined bf16", flush=True)
    errs = []
    for iters in (1, 2, 4):
        Z = [torch.zeros_like(x) for x in X]
        B.solve_refined(X, Z, iters)
        torch.cuda.synchronize()
        errs.append(max(orc_rel(orc_op(orc, T, z), x) for T, x, z in zip(trees, X, Z)))
    assert errs[0] < TOL["bf16"]
    assert errs[1] < errs[0] / 10
    assert errs[2] < TOL["f32"]


def test_batch_fp32_matches_fp64_batch(batch):
    """The FP32 batched solve against the FP64 batched solve of the same
    right-hand sides."""
    import torch
    trees, plans, B, X = batch
    for P in plans:
        assert P.factor() == 0
    Z32 = [torch.zeros_like(x) for x in X]
    B.solve(X, Z32, prec="f32")
    X64 = [x.double() for x in X]
    Z64 = [torch.zeros_like(x) for x in X64]
    B.solve(X64, Z64)
    torch.cuda.synchronize()
    for k, (a, b) in enumerate(zip(Z32, Z64)):
        assert orc_rel(a, b) < TOL["f32"], SPECS[k]


def test_fp32_tuner(batch):
    import torch
    trees, plans, B, X = batch
    Y = [torch.zeros_like(x) for x in X]
    tile, ms = B.tune(X, Y, reps=1, prec="f32")
    assert tile in B.TILES and all(v > 0 for v in ms.values())


def test_grad_outer_f32_matches_f64(one):
    """The FP32 projected outer product against the FP64 one on the same
    (FP32-representable) inputs."""
    import torch
    t, P = one
    g = torch.Generator(device="cpu").manual_seed(8)
    Ga = torch.randn(2, t.m, generator=g).to("cuda")
    Gb = torch.randn(2, t.m, generator=g).to("cuda")
    g32 = P.grad_outer_f32(-0.5, Ga, Gb)
    g64 = P.grad_outer(-0.5, Ga.double(), Gb.double())
    for a, b in zip(g32, g64):
        assert orc_rel(a, b) < 1e-5


def test_fp32_autograd_solve_matches_fp64(one):
    """q = b^T A^{-1} b in FP32 through the autograd wrapper: dq/db and the
    factor gradients against the FP64 path."""
    import torch
    from dplr_amd.autograd import plr_solve
    t, P = one
    assert P.factor() == 0
    b64 = torch.randn(1, t.m, dtype=torch.float64, device="cuda")
    out = {}
    for dt in (torch.float32, torch.float64):
        b = b64.to(dt).clone().requires_grad_(True)
        q = (b * plr_solve(b, P)).sum()
        q.backward()
        out[dt] = (q.item(), b.grad.clone(), [g.clone() for g in P.last_factor_grad])
    q32, gb32, gf32 = out[torch.float32]
    q64, gb64, gf64 = out[torch.float64]
    assert abs(q32 - q64) / abs(q64) < 1e-5
    assert orc_rel(gb32, gb64) < TOL["f32"]
    for a, b in zip(gf32, gf64):
        assert orc_rel(a, b) < 1e-4
