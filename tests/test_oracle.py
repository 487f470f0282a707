"""CPU tests that pin the oracle (oracle/dplr_oracle.c).

Each oracle function is checked against something it does not share code
with: the dense closed form computed by numpy, algebraic identities, the
mathematics of the inverse, and the reference sources themselves compiled
into oracle/_ref (reference src/dplrpmv.c, src/dplrpdmm.c, src/dplrinv.c,
src/dplrprk.c) when that build exists.
"""
import ctypes

import numpy as np
import pytest

import dplr_amd as D
from dplr_amd.tree import PlrTree

PD = ctypes.POINTER(ctypes.c_double)


def rel(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n else 1.0)


SHAPES = [(256, 3, 8), (512, 4, 16), (96, 5, 2), (64, 0, 0), (128, 2, 64)]


@pytest.mark.parametrize("m,nlvl,rank", SHAPES)
@pytest.mark.parametrize("trans", ["N", "T", "c"])
def test_gepmv_closed_form(orc, m, nlvl, rank, trans):
    A = PlrTree.random(m, nlvl, rank, seed=3)
    Ad = A.dense()
    op = Ad if trans == "N" else Ad.T
    rng = np.random.default_rng(1)
    x, y0 = rng.standard_normal(m), rng.standard_normal(m)
    for alpha, beta in [(1.0, 0.0), (0.7, -1.3), (-2.0, 1.0)]:
        y = D.gepmv(trans, A, x, alpha, beta, y0, api=orc)
        assert rel(y, alpha * op @ x + beta * y0) < 1e-13


def test_gepmv_alpha_zero_quirk(orc):
    """alpha == 0 scales y by beta twice (reference src/dplrpmv.c:140-151)."""
    A = PlrTree.random(256, 3, 8, seed=4)
    y0 = np.random.default_rng(2).standard_normal(256)
    y = D.gepmv("N", A, y0, 0.0, 0.5, y0, api=orc)
    assert np.array_equal(y, (y0 * 0.5) * 0.5)
    y = D.gepmv("T", A, y0, 0.0, 0.0, np.full(256, np.nan), api=orc)
    assert np.array_equal(y, np.zeros(256))


def test_gepmv_strides(orc):
    A = PlrTree.random(256, 2, 4, seed=5)
    Ad = A.dense()
    rng = np.random.default_rng(3)
    xs, ys = rng.standard_normal(256 * 2), rng.standard_normal(256 * 3)
    y = D.gepmv("N", A, xs, 1.5, 2.0, ys, incx=2, incy=3, api=orc)
    expect = ys.copy()
    expect[::3] = 1.5 * Ad @ xs[::2] + 2.0 * ys[::3]
    assert rel(y, expect) < 1e-13


def test_adjoint_identity(orc):
    A = PlrTree.random(512, 4, 8, seed=6)
    rng = np.random.default_rng(4)
    x, w = rng.standard_normal(512), rng.standard_normal(512)
    ax = D.gepmv("N", A, x, api=orc)
    atw = D.gepmv("T", A, w, api=orc)
    assert abs(w @ ax - atw @ x) < 1e-12 * np.linalg.norm(w) * np.linalg.norm(ax)


@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
def test_gepdmm_closed_form(orc, ta, tb):
    m, n = 256, 7
    A = PlrTree.random(m, 3, 8, seed=7)
    Ad = A.dense()
    opA = Ad if ta == "N" else Ad.T
    rng = np.random.default_rng(5)
    B = rng.standard_normal((m, n) if tb == "N" else (n, m))
    opB = B if tb == "N" else B.T
    C0 = rng.standard_normal((m, n))
    C = D.gepdmm(ta, tb, A, B, n, 0.5, 2.0, C0, api=orc)
    assert rel(C, 0.5 * opA @ opB + 2.0 * C0) < 1e-13


def test_gepdmm_is_columnwise_gepmv(orc):
    m, n = 512, 5
    A = PlrTree.random(m, 4, 8, seed=8)
    B = np.random.default_rng(6).standard_normal((m, n))
    C = D.gepdmm("N", "N", A, B, n, api=orc)
    for j in range(n):
        assert rel(C[:, j], D.gepmv("N", A, B[:, j], api=orc)) < 1e-14


@pytest.mark.parametrize("m,nlvl,rank", [(256, 3, 8), (512, 4, 16), (128, 1, 8), (64, 0, 0)])
def test_geinv_is_inverse(orc, m, nlvl, rank):
    A = PlrTree.random(m, nlvl, rank, seed=9)
    B, info = D.geinv(A, 0.0, api=orc)
    assert info == 0
    Bd, Ad = B.dense(api=orc), A.dense()
    assert np.abs(Bd @ Ad - np.eye(m)).max() < 1e-12
    assert rel(Bd, np.linalg.inv(Ad)) < 1e-12


def test_geinv_structure(orc):
    """Ranks of inv(A) grow to r*(depth+1); columns are unit norm; sig >= 0."""
    r = 4
    A = PlrTree.random(256, 3, r, seed=10)
    B, info = D.geinv(A, 0.0, api=orc)
    assert info == 0
    for d, n, which, node in B.blocks():
        u, s, v = B.lr_arrays(node)
        assert u.shape[1] == r * (d + 1)
        assert np.all(s >= 0)
        assert np.allclose(np.linalg.norm(u, axis=0), 1.0, atol=1e-13)
        assert np.allclose(np.linalg.norm(v, axis=0), 1.0, atol=1e-13)


def test_geinv_singular_leaf(orc):
    A = PlrTree.random(128, 1, 4, seed=11)
    leaf = A.root.dh.nw.contents
    ctypes.memset(leaf.dd.d, 0, 64 * 64 * 8)
    B, info = D.geinv(A, 0.0, api=orc)
    assert info > 0


def test_geprk_closed_form(orc):
    m, k = 256, 3
    A = PlrTree.random(m, 3, 4, seed=12)
    before = A.dense()
    ranks0 = [A.lr_arrays(nd)[0].shape[1] for (_, _, _, nd) in A.blocks()]
    rng = np.random.default_rng(7)
    u, v, s = rng.standard_normal((m, k)), rng.standard_normal((m, k)), rng.random(k) + 0.1
    D.geprk(A, u, v, s, alpha=-0.5, api=orc)
    assert rel(A.dense(), before - 0.5 * u @ np.diag(s) @ v.T) < 1e-13
    ranks1 = [A.lr_arrays(nd)[0].shape[1] for (_, _, _, nd) in A.blocks()]
    assert ranks1 == [r + k for r in ranks0]


def test_getrf_getri(orc):
    lib = orc["_lib"]
    n = 37
    rng = np.random.default_rng(8)
    A = rng.standard_normal((n, n))
    LU = np.asfortranarray(A.copy())
    piv = (ctypes.c_int * n)()
    info = ctypes.c_int()
    lib.orc_dgetrf(n, n, LU.ctypes.data_as(PD), n, piv, ctypes.byref(info))
    assert info.value == 0
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    PA = A.copy()
    for i in range(n):
        p = piv[i] - 1
        PA[[i, p]] = PA[[p, i]]
    assert np.abs(L @ U - PA).max() < 1e-13
    lib.orc_dgetri(n, LU.ctypes.data_as(PD), n, piv, ctypes.byref(info))
    assert np.abs(LU @ A - np.eye(n)).max() < 1e-12


# ---- against the reference sources compiled into oracle/_ref --------------

@pytest.mark.parametrize("m,nlvl,rank", [(256, 3, 8), (2048, 5, 32), (96, 5, 2)])
def test_matches_reference_apply(orc, ref, m, nlvl, rank):
    A = PlrTree.random(m, nlvl, rank, seed=13)
    rng = np.random.default_rng(9)
    x, y0 = rng.standard_normal(m), rng.standard_normal(m)
    for trans in ("N", "T"):
        for alpha, beta in [(1.0, 0.0), (0.3, 0.9), (0.0, 0.5)]:
            a = D.gepmv(trans, A, x, alpha, beta, y0, api=orc)
            b = D.gepmv(trans, A, x, alpha, beta, y0, api=ref)
            assert rel(a, b) < 1e-13
    B = rng.standard_normal((m, 6))
    for ta in ("N", "T"):
        a = D.gepdmm(ta, "N", A, B, 6, 1.0, 0.0, api=orc)
        b = D.gepdmm(ta, "N", A, B, 6, 1.0, 0.0, api=ref)
        assert rel(a, b) < 1e-13


@pytest.mark.parametrize("eps", [0.0, 1e-8])
def test_matches_reference_inverse(orc, ref, eps):
    A = PlrTree.random(512, 4, 8, seed=14)
    B1, i1 = D.geinv(A, eps, api=orc)
    B2, i2 = D.geinv(A, eps, api=ref)
    assert i1 == i2 == 0
    assert rel(B1.dense(api=orc), B2.dense(api=ref)) < 1e-12
    r1 = [B1.lr_arrays(nd)[0].shape[1] for (_, _, _, nd) in B1.blocks()]
    r2 = [B2.lr_arrays(nd)[0].shape[1] for (_, _, _, nd) in B2.blocks()]
    assert r1 == r2


def test_matches_reference_geprk(orc, ref):
    m, k = 256, 2
    rng = np.random.default_rng(10)
    u, v, s = rng.standard_normal((m, k)), rng.standard_normal((m, k)), rng.random(k)
    A1 = PlrTree.random(m, 3, 4, seed=15)
    A2 = PlrTree.random(m, 3, 4, seed=15)
    D.geprk(A1, u, v, s, alpha=2.0, eps=1e-3, api=orc)
    D.geprk(A2, u, v, s, alpha=2.0, eps=1e-3, api=ref)
    assert rel(A1.dense(api=orc), A2.dense(api=ref)) < 1e-13
