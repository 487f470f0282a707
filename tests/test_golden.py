"""The oracle against the golden fixtures of the reference build (CPU)."""
import numpy as np

import dplr_amd as D
import pytest

from golden_cases import CASES, GOLDEN, kernel_tree, rel, replay


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_oracle_reproduces_golden(orc, golden, case):
    replay(*case, golden, orc, 1e-12)


def test_golden_kernel_fixture_is_consistent(golden):
    """The stored dense-LU solution and the reference's inverse applied to b
    agree (pins the fixture itself)."""
    a, b = golden["kernel/inv_b"], golden["kernel/lu_b"]
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-9
    assert golden["kernel/ranks"].sum() == golden["kernel/sig"].size
    assert np.all(np.diff(golden["kernel/sig"][:golden["kernel/ranks"][0]]) <= 0)


def test_kernel_tree_file_matches_fixture(orc, golden):
    """The stored reference-compressed kernel tree has the ranks and sig the
    fixture recorded, and the oracle's matvec and inverse on it give the
    reference's outputs."""
    T = kernel_tree()
    sig = [T.lr_arrays(nd)[1] for (_, _, _, nd) in T.blocks()]
    assert np.array_equal(np.array([len(s) for s in sig]), golden["kernel/ranks"])
    assert np.array_equal(np.concatenate(sig), golden["kernel/sig"])
    rng = np.random.default_rng(1)
    v, b = rng.standard_normal(T.m), rng.standard_normal(T.m)
    assert rel(D.gepmv("N", T, v, api=orc), golden["kernel/y"]) < 1e-12
    B, info = D.geinv(T, 0.0, api=orc)
    assert info == 0
    assert rel(D.gepmv("N", B, b, api=orc), golden["kernel/inv_b"]) < 1e-9
    B.free()
    T.free()
