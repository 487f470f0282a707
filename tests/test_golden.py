"""The oracle against the golden fixtures of the reference build (CPU)."""
import numpy as np
import pytest

from golden_cases import CASES, GOLDEN, replay


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
