"""Writes tests/golden/golden.npz: outputs of the reference's own C sources
(built into oracle/_ref by oracle/Makefile against the OpenBLAS 0.3.28 in
SciPy) on seeded inputs, so that parity is pinned without the reference.

Inputs are regenerated by the tests from the same seeds: the PLR matrices
by dplrx_gen_random (deterministic C generator of this library), vectors by
numpy's PCG64 (stable across numpy versions).  The input trees of the two
smaller cases, and the Gaussian-kernel tree the reference's dplrd2gep
compressed, are also stored as PLR files (dplrx_save: c1s.plr, deep.plr,
kernel.plr), so the fixtures do not depend on the generator or on an SVD
reproducing them.  Run in the build container:

    python tests/golden/make_golden.py            # golden.npz and the trees
    python tests/golden/make_golden.py --trees    # the PLR files only
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

import dplr_amd as D  # noqa: E402
import oracle  # noqa: E402
from dplr_amd.tree import PlrTree  # noqa: E402

# (name, m, nlvl, rank, seed)
CASES = [("c1s", 1024, 4, 8, 1), ("deep", 512, 5, 4, 2), ("wide", 2048, 2, 48, 3)]
TREE_FILES = ("c1s", "deep")   # the wide case's tree (11 MB) is regenerated only


def gaussian(m, ell, nugget, seed=0):
    x = np.sort(np.random.default_rng(seed).uniform(0, 1, m))
    A = np.exp(-(x[:, None] - x[None, :]) ** 2 / (2 * ell ** 2)) + nugget * np.eye(m)
    return np.asfortranarray(A)


def kernel_tree(ref):
    """N=2048, l=0.1, nugget 1e-2, compressed by the reference's dplrd2gep at
    tol 1e-10 with 5 levels (leaf 64)."""
    m = 2048
    G = gaussian(m, 0.1, 1e-2)
    T = PlrTree(m)
    info = ctypes.c_int(0)
    ref["dplrd2gep"](m, G.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), m, T.ref, 1e-10, 5,
                     ctypes.byref(info))
    assert info.value == 0
    return T


def write_trees(ref):
    for name, m, nlvl, r, seed in CASES:
        if name in TREE_FILES:
            A = PlrTree.random(m, nlvl, r, 4.0, seed)
            A.save(os.path.join(HERE, f"{name}.plr"))
            A.free()
    T = kernel_tree(ref)
    T.save(os.path.join(HERE, "kernel.plr"))
    T.free()


def main():
    ref = oracle.load_ref()
    if ref is None:
        raise SystemExit("oracle/_ref is not built (make -C oracle)")
    write_trees(ref)
    if "--trees" in sys.argv:
        return
    out = {}
    for name, m, nlvl, r, seed in CASES:
        A = PlrTree.random(m, nlvl, r, 4.0, seed)
        rng = np.random.default_rng(seed)
        x = rng.standard_normal(m)
        y0 = rng.standard_normal(m)
        X = np.asfortranarray(rng.standard_normal((m, 4)))
        C0 = np.asfortranarray(rng.standard_normal((m, 4)))
        b = rng.standard_normal(m)
        out[f"{name}/dense_norm"] = np.linalg.norm(A.dense(api=ref))
        out[f"{name}/y_N"] = D.gepmv("N", A, x, 1.5, 0.5, y0.copy(), api=ref)
        out[f"{name}/y_T"] = D.gepmv("T", A, x, -0.5, 2.0, y0.copy(), api=ref)
        out[f"{name}/y_a0"] = D.gepmv("N", A, x, 0.0, 0.5, y0.copy(), api=ref)
        out[f"{name}/C_NN"] = D.gepdmm("N", "N", A, X, 4, 1.0, 0.5, C0.copy(order="F"), api=ref)
        out[f"{name}/C_TN"] = D.gepdmm("T", "N", A, X, 4, 2.0, 0.0, None, api=ref)
        B, info = D.geinv(A, 0.0, api=ref)
        assert info == 0
        out[f"{name}/inv_b"] = D.gepmv("N", B, b, api=ref)
        out[f"{name}/inv_ranks"] = np.array(
            [B.lr_arrays(nd)[0].shape[1] for (_, _, _, nd) in B.blocks()])
        B.free()
        A.free()
    # kernel matrix: N=2048, l=0.1, nugget 1e-2, dplrd2gep tol 1e-10 (leaf 64)
    m = 2048
    T = kernel_tree(ref)
    sig = [T.lr_arrays(nd)[1] for (_, _, _, nd) in T.blocks()]
    out["kernel/ranks"] = np.array([len(s) for s in sig])
    out["kernel/sig"] = np.concatenate(sig)
    rng = np.random.default_rng(1)
    v, b = rng.standard_normal(m), rng.standard_normal(m)
    out["kernel/y"] = D.gepmv("N", T, v, api=ref)
    B, info = D.geinv(T, 0.0, api=ref)
    assert info == 0
    out["kernel/inv_b"] = D.gepmv("N", B, b, api=ref)
    out["kernel/lu_b"] = np.linalg.solve(T.dense(api=ref), b)
    B.free()
    T.free()
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    print(f"wrote {len(out)} arrays")


if __name__ == "__main__":
    main()
