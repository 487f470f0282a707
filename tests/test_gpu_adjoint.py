"""Adjoint kernels (transposed solve/apply and dplrx_grad_outer) through the
autograd wrappers: against dense FP64 reference gradients, against central
finite differences (h = 1e-6, FP64), and on a Gaussian-kernel PLR compressed
by the reference's dplrd2gep (tol 1e-10)."""
import ctypes

import numpy as np
import pytest

import dplr_amd as D
from dplr_amd.abi import PLR_DMTX
from dplr_amd.tree import PlrTree

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def leaf_nodes(T):
    out = []

    def walk(node):
        c = node.contents
        if c.type == PLR_DMTX:
            out.append(node)
            return
        walk(c.dh.nw)
        walk(c.dh.se)
    walk(ctypes.pointer(T.root))
    return out


def block_index(T):
    """(d, q, node) per low-rank block: its U rows are segment q of level d."""
    return [(d, 2 * n if w == "ne" else 2 * n + 1, nd) for (d, n, w, nd) in T.blocks()]


def dense_grads(T, G):
    """Projection of a dense gradient G onto the factors of T, per block."""
    m = T.m
    out = {}
    for (d, q, nd) in block_index(T):
        seg = m >> (d + 1)
        rows = slice(q * seg, (q + 1) * seg)
        cols = slice((q ^ 1) * seg, ((q ^ 1) + 1) * seg)
        u, s, v = T.lr_arrays(nd)
        Gb = G[rows, cols]
        out[(d, q)] = ((Gb @ v) * s, (Gb.T @ u) * s, np.einsum("ik,ij,jk->k", u, Gb, v))
    b = m >> T.depth()
    leaves = [G[l * b:(l + 1) * b, l * b:(l + 1) * b] for l in range(m // b)]
    return out, leaves


def gpu_grads(T, P, grads):
    dU, dV, dS, dD = [g.cpu().numpy() for g in grads]
    col, so = P.layout()
    m = T.m
    out = {}
    for (d, q, nd) in block_index(T):
        seg = m >> (d + 1)
        r = T.lr_arrays(nd)[0].shape[1]
        R = P.ranks[d]
        gu = dU[col[d]:col[d] + r, q * seg:(q + 1) * seg].T
        # the V rows of this block live on segment q^1
        gv = dV[col[d]:col[d] + r, (q ^ 1) * seg:((q ^ 1) + 1) * seg].T
        gs = dS[so[d] + q * R: so[d] + q * R + r]
        out[(d, q)] = (gu, gv, gs)
    leaves = [dD[l].T for l in range(dD.shape[0])]
    return out, leaves


def compare(T, P, grads, G, tol):
    ref, ref_leaves = dense_grads(T, G)
    got, got_leaves = gpu_grads(T, P, grads)
    for key in ref:
        for a, b in zip(got[key], ref[key]):
            if b.size:
                assert rel(a, b) < tol, key
    for a, b in zip(got_leaves, ref_leaves):
        assert rel(a, b) < tol


@pytest.mark.parametrize("m,nlvl,rank,nrhs", [(512, 3, 8, 2), (1024, 4, 12, 1)])
def test_solve_adjoint_matches_dense(cuda, m, nlvl, rank, nrhs):
    import torch
    from dplr_amd.autograd import plr_solve
    print(f"solve adjoint m={m} L={nlvl} r={rank} n={nrhs}", flush=True)
    T = PlrTree.random(m, nlvl, rank, 4.0, 61)
    P = D.Plan(T)
    assert P.factor() == 0
    rng = np.random.default_rng(3)
    b = torch.tensor(rng.standard_normal((nrhs, m)), device=cuda, requires_grad=True)
    w = torch.tensor(rng.standard_normal((nrhs, m)), device=cuda)
    y = plr_solve(b, P)
    (w * y).sum().backward()
    A = T.dense()
    Y = np.linalg.solve(A, b.detach().cpu().numpy().T)    # m x n
    Lam = np.linalg.solve(A.T, w.cpu().numpy().T)          # m x n
    assert rel(y.detach().cpu().numpy().T, Y) < 1e-12
    assert rel(b.grad.cpu().numpy().T, Lam) < 1e-12
    compare(T, P, P.last_factor_grad, -Lam @ Y.T, 1e-10)
    P.close()
    T.free()


def test_apply_adjoint_matches_dense(cuda):
    import torch
    from dplr_amd.autograd import plr_apply
    T = PlrTree.random(1024, 4, 16, 4.0, 62)
    P = D.Plan(T)
    rng = np.random.default_rng(4)
    x = torch.tensor(rng.standard_normal((3, 1024)), device=cuda, requires_grad=True)
    w = torch.tensor(rng.standard_normal((3, 1024)), device=cuda)
    y = plr_apply(x, P)
    (w * y).sum().backward()
    A = T.dense()
    assert rel(x.grad.cpu().numpy().T, A.T @ w.cpu().numpy().T) < 1e-13
    compare(T, P, P.last_factor_grad, w.cpu().numpy().T @ x.detach().cpu().numpy(), 1e-12)
    P.close()
    T.free()


def test_quadratic_form_gradient(cuda):
    """q = b^T A^{-1} b: dq/db = (A^{-1} + A^{-T}) b, dq/dA = -A^{-T} b b^T A^{-T}."""
    import torch
    from dplr_amd.autograd import plr_solve
    T = PlrTree.random(512, 3, 8, 4.0, 63)
    P = D.Plan(T)
    assert P.factor() == 0
    b = torch.tensor(np.random.default_rng(5).standard_normal((1, 512)), device=cuda,
                     requires_grad=True)
    q = (b * plr_solve(b, P)).sum()
    q.backward()
    A = T.dense()
    bb = b.detach().cpu().numpy().ravel()
    assert rel(b.grad.cpu().numpy().ravel(), np.linalg.solve(A, bb) + np.linalg.solve(A.T, bb)) < 1e-12
    y, lam = np.linalg.solve(A, bb), np.linalg.solve(A.T, bb)
    compare(T, P, P.last_factor_grad, -np.outer(lam, y), 1e-10)
    P.close()
    T.free()


def _loss(T, b, w, cuda, refine=False):
    """w . A^{-1} b in FP64; with refine, one step of iterative refinement
    (residual through the FP64 product), which takes the solve's rounding
    error, and so the finite differences' noise, down by the factor of one
    more solve."""
    import torch
    P = D.Plan(T)
    assert P.factor() == 0
    y = torch.zeros_like(b)
    P.solve("N", b, y)
    if refine:
        r = b.clone()
        P.gepdmm("N", -1.0, y, 1.0, r)
        dy = torch.zeros_like(b)
        P.solve("N", r, dy)
        y += dy
    val = float((w * y).sum())
    P.close()
    return val


def test_central_finite_differences(cuda):
    """Single factor entries perturbed in the host tree (h = 1e-6, FP64)."""
    import torch
    from dplr_amd.autograd import plr_solve
    m, h = 512, 1e-6
    T = PlrTree.random(m, 3, 8, 4.0, 64)
    P = D.Plan(T)
    assert P.factor() == 0
    rng = np.random.default_rng(6)
    b = torch.tensor(rng.standard_normal((2, m)), device=cuda)
    w = torch.tensor(rng.standard_normal((2, m)), device=cuda)
    bb = b.clone().requires_grad_(True)
    (w * plr_solve(bb, P)).sum().backward()
    got, got_leaves = gpu_grads(T, P, P.last_factor_grad)
    P.close()
    blocks = {(d, q): nd for (d, q, nd) in block_index(T)}
    checks = [("u", (1, 0), 3, 2), ("v", (2, 7), 5, 1), ("s", (0, 1), None, 4),
              ("u", (0, 1), 200, 0), ("leaf", 2, (1, 7), None)]
    for kind, key, i, k in checks:
        if kind == "leaf":
            node = leaf_nodes(T)[key].contents
            bsz = node.size
            idx = i[1] * bsz + i[0]
            arr, analytic = node.dd.d, got_leaves[key][i[0], i[1]]
        else:
            lr = blocks[key].contents.dlr
            hsz = blocks[key].contents.size
            if kind == "u":
                arr, idx, analytic = lr.u, k * hsz + i, got[key][0][i, k]
            elif kind == "v":
                arr, idx, analytic = lr.v, k * hsz + i, got[key][1][i, k]
            else:
                arr, idx, analytic = lr.sig, k, got[key][2][k]
        x0 = arr[idx]
        arr[idx] = x0 + h
        fp = _loss(T, b, w, cuda)
        arr[idx] = x0 - h
        fm = _loss(T, b, w, cuda)
        arr[idx] = x0
        fd = (fp - fm) / (2 * h)
        print(kind, key, i, k, fd, analytic, flush=True)
        assert abs(fd - analytic) <= 1e-6 * abs(analytic) + 1e-7, (kind, key)
    T.free()


def test_gaussian_kernel_plr_gradient(cuda, ref):
    """N=2048 Gaussian kernel (l = 0.05, nugget 1e-2) compressed by the
    reference's dplrd2gep at tol 1e-10 (leaf 128): dense gradient check."""
    import torch
    from dplr_amd.autograd import plr_solve
    m, L = 2048, 4
    x = np.sort(np.random.default_rng(0).uniform(0, 1, m))
    A = np.exp(-(x[:, None] - x[None, :]) ** 2 / (2 * 0.05 ** 2)) + 1e-2 * np.eye(m)
    A = np.asfortranarray(A)
    T = PlrTree(m)
    info = ctypes.c_int(0)
    ref["dplrd2gep"](m, A.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), m, T.ref, 1e-10, L,
                     ctypes.byref(info))
    assert info.value == 0
    Ad = T.dense(api=ref)
    assert rel(Ad, A) < 1e-8
    P = D.Plan(T)
    assert P.factor() == 0
    b = torch.tensor(np.random.default_rng(1).standard_normal((1, m)), device=cuda,
                     requires_grad=True)
    q = (b * plr_solve(b, P)).sum()
    q.backward()
    bb = b.detach().cpu().numpy().ravel()
    y, lam = np.linalg.solve(Ad, bb), np.linalg.solve(Ad.T, bb)
    assert rel(b.grad.cpu().numpy().ravel(), y + lam) < 1e-9
    compare(T, P, P.last_factor_grad, -np.outer(lam, y), 1e-8)
    P.close()


def test_gaussian_kernel_gradient_workload_size(cuda, orc):
    """The gradient workload at its stated size: N=16384 Gaussian kernel
    (x ~ U[0,1] sorted, l = 0.05, nugget 1e-2) compressed at tol 1e-10 with
    leaf 128, q = b^T A^{-1} b.  FP64: y against the oracle's product
    (residual), dq/db = 2y (A is symmetric), central finite differences
    (h = 1e-6) on factor entries of three levels and a leaf.  FP32: q,
    dq/db and the factor gradients against FP64 (tolerances: the condition
    number, ~1e4 here, times FP32 rounding)."""
    import torch
    from dplr_amd.autograd import plr_solve
    m, L, h = 16384, 7, 1e-6
    x = np.sort(np.random.default_rng(0).uniform(0, 1, m))
    A = np.asfortranarray(np.exp(-(x[:, None] - x[None, :]) ** 2 / (2 * 0.05 ** 2)) + 1e-2 * np.eye(m))
    T = PlrTree(m)
    info = ctypes.c_int(0)
    D.lib.dplrd2gep(m, A.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), m, T.ref, 1e-10, L,
                    ctypes.byref(info))
    assert info.value == 0
    del A
    P = D.Plan(T)
    assert P.factor() == 0
    b = torch.tensor(np.random.default_rng(1).standard_normal((1, m)), device=cuda)
    out = {}
    for dt in (torch.float64, torch.float32):
        if dt == torch.float32:
            assert P.factor_single("f32") == 0
        bb = b.to(dt).clone().requires_grad_(True)
        y = plr_solve(bb, P)
        q = (bb * y).sum()
        q.backward()
        out[dt] = (q.item(), y.detach().clone(), bb.grad.clone(),
                   [g.clone() for g in P.last_factor_grad])
    q64, y64, g64, f64 = out[torch.float64]
    yh = y64.cpu().numpy().ravel()
    back = D.gepdmm("N", "N", T, yh.reshape(m, 1).copy(), 1, api=orc).ravel()
    # condition number ~2e5 (row sums ~2e3, nugget 1e-2); the factored solve's
    # residual measured 1.7e-9
    res = rel(back, b.cpu().numpy().ravel())
    print(f"gradient workload: FP64 solve relative residual {res:.2e}", flush=True)
    assert res < 2e-8
    assert rel(g64.cpu().numpy(), 2 * yh) < 1e-9
    got, got_leaves = gpu_grads(T, P, f64)
    blocks = {(d, qq): nd for (d, qq, nd) in block_index(T)}
    checks = [("u", (0, 1), 100, 0), ("v", (3, 5), 7, 1), ("s", (6, 40), None, 0),
              ("leaf", 17, (3, 9), None)]
    w = b.clone()
    for kind, key, i, k in checks:
        if kind == "leaf":
            node = leaf_nodes(T)[key].contents
            idx = i[1] * node.size + i[0]
            arr, analytic = node.dd.d, got_leaves[key][i[0], i[1]]
        else:
            lr = blocks[key].contents.dlr
            hsz = blocks[key].contents.size
            if kind == "u":
                arr, idx, analytic = lr.u, k * hsz + i, got[key][0][i, k]
            elif kind == "v":
                arr, idx, analytic = lr.v, k * hsz + i, got[key][1][i, k]
            else:
                arr, idx, analytic = lr.sig, k, got[key][2][k]
        x0 = arr[idx]
        arr[idx] = x0 + h
        fp = _loss(T, b, w, cuda, refine=True)
        arr[idx] = x0 - h
        fm = _loss(T, b, w, cuda, refine=True)
        arr[idx] = x0
        fd = (fp - fm) / (2 * h)
        print("fd", kind, key, fd, analytic, flush=True)
        # the difference quotient's error here: rounding of q over 2h and the
        # h^2 term at a condition number of ~2e5 (a deep level's sig is ~1e-13,
        # far below h); each loss is refined once, as a single factored solve's
        # rounding of q (relative residual ~1e-9) alone moves the quotient by ~1
        assert abs(fd - analytic) <= 1e-4 * abs(analytic) + 1e-3, (kind, key)
    q32, y32, g32, f32 = out[torch.float32]
    assert abs(q32 - q64) / abs(q64) < 1e-5
    assert rel(g32.double().cpu().numpy(), g64.cpu().numpy()) < 2e-3
    for a, bref in zip(f32, f64):
        assert rel(a.double().cpu().numpy(), bref.cpu().numpy()) < 5e-2
    P.close()
    T.free()
