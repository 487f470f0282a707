"""Shared replay of the golden fixtures (tests/golden/golden.npz, written by
tests/golden/make_golden.py from the reference build) on any api: the oracle
on the CPU, the product on the GPU."""
import os

import numpy as np

import dplr_amd as D
from dplr_amd.tree import PlrTree

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GOLDEN = os.path.join(GOLDEN_DIR, "golden.npz")
CASES = [("c1s", 1024, 4, 8, 1), ("deep", 512, 5, 4, 2), ("wide", 2048, 2, 48, 3)]


def case_tree(name, m, nlvl, r, seed):
    """The input tree of a case: from its PLR file when the fixtures hold one
    (checked to equal what the generator makes), else from the generator."""
    A = PlrTree.random(m, nlvl, r, 4.0, seed)
    path = os.path.join(GOLDEN_DIR, f"{name}.plr")
    if not os.path.exists(path):
        return A
    F = PlrTree.load(path)
    assert F.m == m and np.array_equal(F.dense(), A.dense()), name
    A.free()
    return F


def kernel_tree():
    """The N=2048 Gaussian-kernel tree compressed by the reference's
    dplrd2gep (tests/golden/kernel.plr)."""
    return PlrTree.load(os.path.join(GOLDEN_DIR, "kernel.plr"))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def replay(name, m, nlvl, r, seed, g, api, tol):
    """Every fixture of one case through `api` (None = product)."""
    A = case_tree(name, m, nlvl, r, seed)
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(m)
    y0 = rng.standard_normal(m)
    X = np.asfortranarray(rng.standard_normal((m, 4)))
    C0 = np.asfortranarray(rng.standard_normal((m, 4)))
    b = rng.standard_normal(m)
    assert abs(np.linalg.norm(A.dense()) - g[f"{name}/dense_norm"]) < 1e-12 * g[f"{name}/dense_norm"]
    assert rel(D.gepmv("N", A, x, 1.5, 0.5, y0.copy(), api=api), g[f"{name}/y_N"]) < tol
    assert rel(D.gepmv("T", A, x, -0.5, 2.0, y0.copy(), api=api), g[f"{name}/y_T"]) < tol
    assert rel(D.gepmv("N", A, x, 0.0, 0.5, y0.copy(), api=api), g[f"{name}/y_a0"]) < tol
    assert rel(D.gepdmm("N", "N", A, X, 4, 1.0, 0.5, C0.copy(order="F"), api=api),
               g[f"{name}/C_NN"]) < tol
    assert rel(D.gepdmm("T", "N", A, X, 4, 2.0, 0.0, None, api=api), g[f"{name}/C_TN"]) < tol
    B, info = D.geinv(A, 0.0, api=api)
    assert info == 0
    ranks = [B.lr_arrays(nd)[0].shape[1] for (_, _, _, nd) in B.blocks()]
    assert np.array_equal(np.array(ranks), g[f"{name}/inv_ranks"])
    assert rel(D.gepmv("N", B, b, api=api), g[f"{name}/inv_b"]) < tol
    B.free()
    A.free()
