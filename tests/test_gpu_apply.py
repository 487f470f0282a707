"""GPU parity of the apply path (dplrgepmv, dplrgepdmm and the dplrx plan API)
against the oracle on identical inputs; tolerance 1e-12 relative (2-norm)."""
import ctypes

import numpy as np
import pytest

import dplr_amd as D
from dplr_amd.tree import PlrTree

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n else 1.0)


SHAPES = [
    (2048, 5, 32),   # leaf 64: uniform-segment fast path
    (2048, 6, 16),   # leaf 32: sub-wavefront segments
    (1024, 2, 40),   # rank not a multiple of the rank group
    (96, 5, 2),      # leaf 3, non power-of-two segments
    (64, 0, 0),      # dense root
    (256, 3, 0),     # all ranks zero
    (6144, 5, 16),   # leaf 192: segments of 192 rows, not a power-of-two leaf
    (4096, 4, 24),   # leaf 256: one leaf per 256-row expansion group
    (8192, 3, 16),   # leaf 1024, the largest leaf of the two-launch matvec
    (128, 1, 4),     # N = 128: 128-row expansion groups
    (16384, 5, 20),  # leaf 512, segments longer than 1024 rows (partials)
]


@pytest.mark.parametrize("m,nlvl,rank", SHAPES)
@pytest.mark.parametrize("trans", ["N", "T"])
def test_gepmv_dropin(cuda, orc, m, nlvl, rank, trans):
    print(f"gepmv m={m} nlvl={nlvl} rank={rank} trans={trans}", flush=True)
    A = PlrTree.random(m, nlvl, rank, seed=21)
    rng = np.random.default_rng(1)
    x, y0 = rng.standard_normal(m), rng.standard_normal(m)
    for alpha, beta in [(1.0, 0.0), (0.7, -1.3), (-2.0, 1.0)]:
        y = D.gepmv(trans, A, x, alpha, beta, y0)
        yo = D.gepmv(trans, A, x, alpha, beta, y0, api=orc)
        assert rel(y, yo) < TOL, (alpha, beta)


def test_gepmv_quirks_and_strides(cuda, orc):
    A = PlrTree.random(512, 3, 8, seed=22)
    rng = np.random.default_rng(2)
    y0 = rng.standard_normal(512)
    # alpha == 0 -> beta^2 * y, exactly as the reference
    y = D.gepmv("N", A, y0, 0.0, 0.5, y0)
    assert np.array_equal(y, D.gepmv("N", A, y0, 0.0, 0.5, y0, api=orc))
    # beta == 0 never reads y (NaN in, finite out)
    x = rng.standard_normal(512)
    y = D.gepmv("T", A, x, 1.0, 0.0, np.full(512, np.nan))
    assert np.all(np.isfinite(y))
    xs, ys = rng.standard_normal(1024), rng.standard_normal(512 * 3)
    a = D.gepmv("N", A, xs, 1.5, 2.0, ys, incx=2, incy=3)
    b = D.gepmv("N", A, xs, 1.5, 2.0, ys, incx=2, incy=3, api=orc)
    assert rel(a, b) < TOL


def test_gepmv_varying_ranks(cuda, orc):
    rng = np.random.default_rng(3)
    A = PlrTree.from_blocks(1024, 4, lambda d, n, w: (3 * d + n + (w == "sw") * 5) % 13, rng)
    x = rng.standard_normal(1024)
    for trans in ("N", "T"):
        assert rel(D.gepmv(trans, A, x), D.gepmv(trans, A, x, api=orc)) < TOL


@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
@pytest.mark.parametrize("n", [1, 5, 64])
def test_gepdmm_dropin(cuda, orc, ta, tb, n):
    print(f"gepdmm {ta}{tb} n={n}", flush=True)
    m = 1024
    A = PlrTree.random(m, 4, 24, seed=23)
    rng = np.random.default_rng(4)
    B = rng.standard_normal((m + 3, n) if tb == "N" else (n + 2, m))
    C0 = rng.standard_normal((m + 5, n))
    ldb = B.shape[0]
    a = D.gepdmm(ta, tb, A, B, n, 0.5, 2.0, C0, ldb=ldb, ldc=m + 5)
    b = D.gepdmm(ta, tb, A, B, n, 0.5, 2.0, C0, ldb=ldb, ldc=m + 5, api=orc)
    assert rel(a, b) < TOL


def test_plan_device_api(cuda, orc):
    import torch
    m = 4096
    A = PlrTree.random(m, 6, 32, seed=24)
    P = D.Plan(A)
    assert P.ranks == [32] * 6 and P.leaf == 64
    rng = np.random.default_rng(5)
    x = rng.standard_normal(m)
    xd = torch.tensor(x, device=cuda)
    yd = torch.zeros(m, dtype=torch.float64, device=cuda)
    for trans in ("N", "T"):
        P.gepmv(trans, 1.0, xd, 0.0, yd)
        torch.cuda.synchronize()
        assert rel(yd.cpu().numpy(), D.gepmv(trans, A, x, api=orc)) < TOL
    B = rng.standard_normal((m, 16))
    Bd = torch.tensor(B.T.copy(), device=cuda)      # (n, m) row-major == m x n column-major
    Cd = torch.zeros_like(Bd)
    P.gepdmm("N", 1.0, Bd, 0.0, Cd)
    torch.cuda.synchronize()
    assert rel(Cd.cpu().numpy().T, D.gepdmm("N", "N", A, B, 16, api=orc)) < TOL
    P.close()


def test_full_size_matvec(cuda, orc):
    """N = 65536, 10 levels, leaf 64, rank 32: one RHS each way."""
    m = 65536
    print("full-size matvec", flush=True)
    A = PlrTree.random(m, 10, 32, seed=1)
    x = np.random.default_rng(6).standard_normal(m)
    for trans in ("N", "T"):
        assert rel(D.gepmv(trans, A, x), D.gepmv(trans, A, x, api=orc)) < TOL


@pytest.mark.parametrize("m,nlvl,rank,n", [
    (4096, 6, 32, 70),    # leaf 64, two column groups
    (4096, 5, 24, 16),    # leaf 128
    (2048, 3, 40, 33),    # leaf 256, rank not a multiple of 16
    (8192, 7, 20, 64),
    (8192, 7, 64, 40),    # rank 64: 28 rank tiles, two passes over the staged X
    (2048, 6, 16, 8),     # leaf 32: the general path
    (3072, 5, 16, 8),     # leaf 96: the general path
    (6144, 5, 16, 8),     # leaf 192: segments of 192 rows
    (256, 2, 8, 5),       # no level spans a 256-row group: nothing staged
    (512, 3, 12, 7),      # one staged level of 12 ranks
])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_multi_rhs_product_fused_path(cuda, orc, m, nlvl, rank, n, trans):
    """Even leading dimensions and N a multiple of 256: the all-level
    reduction, rank-major coefficients and the barrier-free expansion."""
    import torch
    print(f"fused product m={m} L={nlvl} r={rank} n={n} {trans}", flush=True)
    A = PlrTree.random(m, nlvl, rank, seed=25)
    P = D.Plan(A)
    rng = np.random.default_rng(7)
    B = rng.standard_normal((m, n))
    C0 = rng.standard_normal((m, n))
    Bd = torch.tensor(B.T.copy(), device=cuda)
    Cd = torch.tensor(C0.T.copy(), device=cuda)
    P.gepdmm(trans, 0.75, Bd, -1.5, Cd)
    torch.cuda.synchronize()
    ref = D.gepdmm(trans, "N", A, B, n, 0.75, -1.5, C0, api=orc)
    assert rel(Cd.cpu().numpy().T, ref) < TOL
    P.close()
