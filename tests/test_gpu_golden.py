"""The GPU product against the golden fixtures of the reference build."""
import ctypes

import numpy as np
import pytest

import dplr_amd as D
from dplr_amd.tree import PlrTree
from golden_cases import CASES, GOLDEN, kernel_tree, rel, replay

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_product_reproduces_golden(cuda, golden, case):
    print("golden", case, flush=True)
    replay(*case, golden, None, 1e-12)


def test_kernel_matrix_golden(cuda, golden):
    """dplrd2gep (GPU SVDs), matvec and inverse applied on the N=2048 kernel
    matrix against the reference's outputs."""
    m = 2048
    x = np.sort(np.random.default_rng(0).uniform(0, 1, m))
    G = np.asfortranarray(np.exp(-(x[:, None] - x[None, :]) ** 2 / (2 * 0.1 ** 2)) + 1e-2 * np.eye(m))
    T = PlrTree(m)
    info = ctypes.c_int(0)
    D.lib.dplrd2gep(m, G.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), m, T.ref, 1e-10, 5,
                    ctypes.byref(info))
    assert info.value == 0
    sig = [T.lr_arrays(nd)[1] for (_, _, _, nd) in T.blocks()]
    ranks = np.array([len(s) for s in sig])
    gr = golden["kernel/ranks"]
    assert np.all(np.abs(ranks - gr) <= 1) and np.sum(ranks != gr) <= max(1, len(gr) // 20)
    off = 0
    for s, k in zip(sig, gr):
        n = min(len(s), k)
        assert np.allclose(s[:n], golden["kernel/sig"][off:off + n], rtol=1e-9, atol=1e-13)
        off += k
    rng = np.random.default_rng(1)
    v, b = rng.standard_normal(m), rng.standard_normal(m)
    assert rel(D.gepmv("N", T, v), golden["kernel/y"]) < 1e-10
    B, info = D.geinv(T, 0.0)
    assert info == 0
    kappa = 5e4
    assert rel(D.gepmv("N", B, b), golden["kernel/inv_b"]) < 50 * kappa * 1e-16 + 1e-9
    B.free()
    T.free()


def test_kernel_tree_file_golden(cuda, golden):
    """The reference-compressed kernel tree (tests/golden/kernel.plr): the GPU
    matvec and the GPU explicit inverse applied to b against the reference's
    outputs on exactly that tree."""
    T = kernel_tree()
    rng = np.random.default_rng(1)
    v, b = rng.standard_normal(T.m), rng.standard_normal(T.m)
    assert rel(D.gepmv("N", T, v), golden["kernel/y"]) < 1e-12
    B, info = D.geinv(T, 0.0)
    assert info == 0
    assert rel(D.gepmv("N", B, b), golden["kernel/inv_b"]) < 1e-9
    B.free()
    T.free()
