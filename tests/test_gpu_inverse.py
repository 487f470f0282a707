"""GPU drop-in dplrgeinv (explicit inverse in the reference's structure) and
dplrgeprk against the oracle and the reference build on identical inputs."""
import ctypes

import numpy as np
import pytest

import dplr_amd as D
from dplr_amd.abi import PLR_HMTX, PLR_LRMTX
from dplr_amd.tree import PlrTree, _fn, nodes

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def ranks(T):
    return [T.lr_arrays(nd)[0].shape[1] for (_, _, _, nd) in T.blocks()]


SHAPES = [(512, 3, 16), (1024, 4, 8), (256, 5, 4), (96, 5, 2), (64, 0, 0), (256, 2, 0)]


@pytest.mark.parametrize("m,nlvl,rank", SHAPES)
def test_geinv_matches_oracle(cuda, orc, m, nlvl, rank):
    print(f"geinv m={m} nlvl={nlvl} rank={rank}", flush=True)
    A = PlrTree.random(m, nlvl, rank, seed=51)
    B, info = D.geinv(A, 0.0)
    Bo, info_o = D.geinv(A, 0.0, api=orc)
    assert info == info_o == 0
    assert ranks(B) == ranks(Bo)
    assert rel(B.dense(), Bo.dense(api=orc)) < 1e-12
    for (_, _, _, nd) in B.blocks():
        u, s, v = B.lr_arrays(nd)
        assert np.all(s >= 0)
        if u.shape[1]:
            assert np.allclose(np.linalg.norm(u, axis=0), 1, atol=1e-12)
            assert np.allclose(np.linalg.norm(v, axis=0), 1, atol=1e-12)
    # inverse applied to a right-hand side, through the oracle's matvec on each B
    b = np.random.default_rng(1).standard_normal(m)
    assert rel(D.gepmv("N", B, b, api=orc), D.gepmv("N", Bo, b, api=orc)) < 1e-12


def test_geinv_rank0_sibling_is_nan_like_reference(cuda, ref):
    """eps = 0 and one off-diagonal block of rank 0: the reference divides a
    zero correction column by its zero norm (src/dplrinv.c:378-404) and the
    NaN spreads through the root's Woodbury step to the whole matrix; the
    drop-in reproduces that (the factored solve does not have the problem,
    see test_gpu_solve.py)."""
    rng = np.random.default_rng(2)
    A = PlrTree.from_blocks(512, 3, lambda d, n, w: 0 if (d, n, w) == (1, 1, "sw") else 4, rng)
    B, info = D.geinv(A, 0.0)
    Br, info_r = D.geinv(A, 0.0, api=ref)
    assert info == info_r == 0
    assert np.array_equal(np.isnan(B.dense()), np.isnan(Br.dense(api=ref)))
    assert np.isnan(Br.dense(api=ref)).any()


def test_geinv_varying_ranks_and_eps(cuda, orc, ref):
    rng = np.random.default_rng(2)
    A = PlrTree.from_blocks(1024, 4, lambda d, n, w: 1 + (3 * d + n + (w == "sw") * 4) % 9, rng)
    B, info = D.geinv(A, 0.0)
    Br, info_r = D.geinv(A, 0.0, api=ref)
    assert info == info_r == 0 and ranks(B) == ranks(Br)
    assert rel(B.dense(), Br.dense(api=ref)) < 1e-12
    # eps > 0: sig below the per-level share of eps is zeroed (reference rule).
    # The reference builds each level from the already truncated subtree
    # inverses, the GPU from the exact factored inverse, so the two differ by
    # a small fraction of the truncation error itself (1.5e-5 here).
    B1, _ = D.geinv(A, 1e-3)
    B2, _ = D.geinv(A, 1e-3, api=ref)
    assert ranks(B1) == ranks(B2)
    z1 = sum(int(np.sum(B1.lr_arrays(nd)[1] == 0)) for (_, _, _, nd) in B1.blocks())
    z2 = sum(int(np.sum(B2.lr_arrays(nd)[1] == 0)) for (_, _, _, nd) in B2.blocks())
    assert z1 == z2 and z1 > 0
    Ai = np.linalg.inv(A.dense(api=ref))
    e_gpu, e_ref = rel(B1.dense(), Ai), rel(B2.dense(api=ref), Ai)
    assert 1e-7 < e_ref < 1e-4
    assert e_gpu <= 1.1 * e_ref
    assert rel(B1.dense(), B2.dense(api=ref)) < 0.05 * e_ref


def test_geinv_reuses_existing_b(cuda, orc):
    A = PlrTree.random(512, 3, 8, seed=52)
    B, _ = D.geinv(A, 0.0)
    info = ctypes.c_int(99)
    D.lib.dplrgeinv(512, A.ref, B.ref, 0.0, ctypes.byref(info))   # second call into the same B
    Bo, _ = D.geinv(A, 0.0, api=orc)
    assert info.value == 0 and rel(B.dense(), Bo.dense(api=orc)) < 1e-12


def _lr_buffers(tree):
    """(u, v, sig) of every low-rank block over its whole capacity, in
    pre-order."""
    out = []

    def walk(p):
        c = p.contents
        if c.type == PLR_LRMTX:
            n, cap = c.size * c.dlr.capacity, c.dlr.capacity
            get = (lambda q, k: np.ctypeslib.as_array(q, shape=(k,)).copy() if k else np.zeros(0))
            out.append((get(c.dlr.u, n), get(c.dlr.v, n), get(c.dlr.sig, cap)))
        elif c.type == PLR_HMTX:
            for w in ("nw", "ne", "se", "sw"):
                walk(getattr(c.dh, w))
    walk(ctypes.pointer(tree.root))
    return out


@pytest.mark.parametrize("reuse", [False, True])
def test_geinv_output_buffers_match_reference(cuda, ref, reuse):
    """The output tree as the reference shapes it: every node's type, size,
    rank and capacity (capacities at least double as correction columns are
    appended, src/dplrresizelr.c:54), and NaN exactly where the reference's
    buffers hold NaN (the slots above the rank of the buffers the call
    allocates) -- for a new B, and for a B shaped by an earlier call."""
    A = PlrTree.from_blocks(1024, 4, lambda d, n, w: 1 + (3 * d + n + (w == "sw") * 4) % 9,
                            np.random.default_rng(5))
    B, info = D.geinv(A, 0.0)
    Br, info_r = D.geinv(A, 0.0, api=ref)
    if reuse:
        i1, i2 = ctypes.c_int(9), ctypes.c_int(9)
        D.lib.dplrgeinv(1024, A.ref, B.ref, 0.0, ctypes.byref(i1))
        _fn(ref, "dplrgeinv")(1024, A.ref, Br.ref, 0.0, ctypes.byref(i2))
        info, info_r = i1.value, i2.value
    assert info == info_r == 0
    assert nodes(B) == nodes(Br)
    for (u, v, s), (ur, vr, sr) in zip(_lr_buffers(B), _lr_buffers(Br)):
        for x, y in ((u, ur), (v, vr), (s, sr)):
            assert np.array_equal(np.isnan(x), np.isnan(y))
    assert rel(B.dense(), Br.dense(api=ref)) < 1e-12


def test_geinv_singular(cuda, orc):
    A = PlrTree.random(256, 2, 4, seed=53)
    leaf = A.root.dh.nw.contents.dh.se.contents
    ctypes.memset(leaf.dd.d, 0, 64 * 64 * 8)
    _, info = D.geinv(A, 0.0)
    _, info_o = D.geinv(A, 0.0, api=orc)
    assert info > 0 and info == info_o


@pytest.mark.parametrize("alpha,eps", [(1.5, 0.0), (-0.7, 0.0), (2.0, 1e-2)])
def test_geprk_matches_reference(cuda, ref, alpha, eps):
    m, k = 512, 3
    rng = np.random.default_rng(3)
    u, v, s = rng.standard_normal((m, k)), rng.standard_normal((m, k)), rng.random(k) + 0.05
    s[1] = 1e-9  # below the thresholds for eps > 0
    A1 = PlrTree.random(m, 3, 4, seed=54)
    A2 = PlrTree.random(m, 3, 4, seed=54)
    D.geprk(A1, u, v, s, alpha=alpha, eps=eps)
    D.geprk(A2, u, v, s, alpha=alpha, eps=eps, api=ref)
    assert ranks(A1) == ranks(A2)
    assert rel(A1.dense(), A2.dense(api=ref)) < 1e-13
    for (_, _, _, n1), (_, _, _, n2) in zip(A1.blocks(), A2.blocks()):
        s1, s2 = A1.lr_arrays(n1)[1], A2.lr_arrays(n2)[1]
        assert np.array_equal(s1 == 0, s2 == 0)
        assert np.allclose(s1, s2, rtol=1e-13, atol=0)
