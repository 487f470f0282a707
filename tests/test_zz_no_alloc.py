"""dplrx_gepmv_noalloc (the matvec whose kernels keep their scratch in the
library's static device array) against dplrx_gepmv, through the C-ABI:
the same kernels run on the same inputs, so the results must be identical.
Two calls in a row, the second on a smaller matrix; both directions, a
nonzero beta, the general-shape kernels (leaf 96, arrival counters in the
scratch), and the size limit."""
import ctypes

import numpy as np
import pytest

import dplr_amd as D
from dplr_amd._lib import lib
from dplr_amd.tree import PlrTree

pytestmark = pytest.mark.gpu


def call(fn, P, trans, alpha, x, beta, y, torch):
    rc = fn(P.p, trans.encode(), alpha, ctypes.c_void_p(x.data_ptr()), beta,
            ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return rc


@pytest.mark.parametrize("trans,beta", [("N", 0.0), ("T", -0.5)])
def test_noalloc_matvec_matches_gepmv(cuda, trans, beta):
    import torch
    specs = [(65536, 10, 32), (4096, 6, 16), (3072, 5, 16)]  # larger, smaller, leaf 96
    rng = np.random.default_rng(12)
    for k, (m, L, r) in enumerate(specs):
        print(f"no-alloc matvec m={m} L={L} r={r} trans={trans}", flush=True)
        A = PlrTree.random(m, L, r, 4.0, seed=60 + k)
        P = D.Plan(A)
        x = torch.tensor(rng.standard_normal(m), device=cuda)
        y0 = torch.tensor(rng.standard_normal(m), device=cuda)
        y_ref, y_new = y0.clone(), y0.clone()
        assert call(lib.dplrx_gepmv, P, trans, 1.25, x, beta, y_ref, torch) == 0
        assert call(lib.dplrx_gepmv_noalloc, P, trans, 1.25, x, beta, y_new, torch) == 0
        torch.cuda.synchronize()
        assert torch.equal(y_new, y_ref), (m, trans)
        # again on the same matrix: the scratch holds the last call's values
        y_new2 = y0.clone()
        assert call(lib.dplrx_gepmv_noalloc, P, trans, 1.25, x, beta, y_new2, torch) == 0
        torch.cuda.synchronize()
        assert torch.equal(y_new2, y_ref), (m, trans)
        P.close()
        A.free()


def test_noalloc_matvec_rejects_large_plans(cuda):
    """N = 262144 on 12 levels (leaf 64) with rank 64 needs 5.1 MB of
    scratch (4.2 MB of coefficients, one per segment and rank, and 0.9 MB of
    partials), more than the 4 MB the entry point holds: a clear error and
    nothing launched (y unchanged)."""
    import torch
    m = 262144
    A = PlrTree.random(m, 12, 64, 4.0, seed=70)
    P = D.Plan(A)
    A.free()
    x = torch.ones(m, dtype=torch.float64, device=cuda)
    y = torch.full((m,), 2.0, dtype=torch.float64, device=cuda)
    rc = call(lib.dplrx_gepmv_noalloc, P, "N", 1.0, x, 0.0, y, torch)
    torch.cuda.synchronize()
    assert rc == -7
    assert b"scratch" in lib.dplrx_last_error()
    assert torch.all(y == 2.0)
    P.close()
