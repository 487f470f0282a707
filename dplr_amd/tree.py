"""Host-side PLR trees (the DPlrMtx layout of include/dplr.h) for Python callers.

A `PlrTree` owns a root node; its buffers and children are malloc'd by the C
library, so any of the C entry points (the product library, the oracle, the
reference build) can read and reshape it.
"""
import ctypes
from ctypes import byref, c_char, c_double, c_int

import numpy as np

from ._lib import PLR_DMTX, PLR_HMTX, PLR_LRMTX, DPlrMtx, lib

_PD = ctypes.POINTER(c_double)


def _fn(api, name):
    """Function `name` from an api dict (oracle, reference build) or the product library."""
    if api is not None:
        return api[name]
    return getattr(lib, name)


def _ptr(a):
    return a.ctypes.data_as(_PD)


class PlrTree:
    """An m x m PLR matrix in host memory."""

    def __init__(self, m=0):
        self.root = DPlrMtx()
        lib.dplrini(byref(self.root))
        self.m = m

    @property
    def ref(self):
        return byref(self.root)

    def free(self):
        if self.root is not None:
            lib.dplrclr(byref(self.root))
            self.root = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # -- construction ------------------------------------------------------
    @classmethod
    def load(cls, path):
        """A tree read from a PLR file (dplrx_load)."""
        t = cls(0)
        m = c_int(0)
        if lib.dplrx_load(str(path).encode(), t.ref, byref(m)) != 0:
            raise RuntimeError(f"dplrx_load failed for {path}")
        t.m = m.value
        return t

    def save(self, path):
        """Write the tree to a PLR file (dplrx_save)."""
        if lib.dplrx_save(str(path).encode(), self.m, self.ref) != 0:
            raise RuntimeError(f"dplrx_save failed for {path}")

    @classmethod
    def random(cls, m, nlvl, rank, shift=4.0, seed=0):
        """Seeded generator of the uniform random PLR family (dplrx_gen_random)."""
        t = cls(m)
        rc = lib.dplrx_gen_random(m, nlvl, rank, shift, seed, t.ref)
        if rc != 0:
            raise RuntimeError("dplrx_gen_random failed")
        return t

    @classmethod
    def from_blocks(cls, m, nlvl, rank_of, rng, shift=4.0, sig_of=None):
        """Tree with arbitrary per-block ranks: rank_of(depth, node, which) with
        which in ('ne', 'sw'); factors are orthonormalised N(0,1) columns."""
        t = cls(m)

        def fill(node, size, d, n):
            if d == nlvl:
                lib.dplrinid(size, node)
                blk = np.asfortranarray(
                    rng.standard_normal((size, size)) / np.sqrt(max(size, 1)) + shift * np.eye(size))
                ctypes.memmove(node.contents.dd.d, blk.ctypes.data, blk.nbytes)
                return
            lib.dplrinih(size, node)
            h = size // 2
            for which in ("ne", "sw"):
                r = min(rank_of(d, n, which), h)
                child = getattr(node.contents.dh, which)
                lib.dplrinilr(h, child, r)
                if r > 0:
                    # keep every array alive while its address is in use
                    u = np.asfortranarray(np.linalg.qr(rng.standard_normal((h, r)))[0])
                    v = np.asfortranarray(np.linalg.qr(rng.standard_normal((h, r)))[0])
                    s = np.ascontiguousarray(
                        sig_of(d, n, which, r) if sig_of else 10.0 ** (-np.arange(r) / 4.0),
                        dtype=np.float64)
                    lr = child.contents.dlr
                    ctypes.memmove(lr.u, u.ctypes.data, u.nbytes)
                    ctypes.memmove(lr.v, v.ctypes.data, v.nbytes)
                    ctypes.memmove(lr.sig, s.ctypes.data, 8 * r)
                    lr.rank = r
            fill(node.contents.dh.nw, h, d + 1, 2 * n)
            fill(node.contents.dh.se, h, d + 1, 2 * n + 1)

        fill(ctypes.pointer(t.root), m, 0, 0)
        return t

    # -- inspection ----------------------------------------------------------
    def dense(self, api=None):
        """Dense expansion via dplrgep2d (api: a dict of C functions; default product)."""
        B = np.zeros((self.m, self.m), order="F")
        f = _fn(api, "dplrgep2d")
        f(self.m, self.ref, _ptr(B), self.m)
        return B

    def blocks(self):
        """Yield (depth, node, which, node_ptr) for every low-rank block."""
        def walk(node, d, n):
            c = node.contents
            if c.type != PLR_HMTX:
                return
            yield (d, n, "ne", c.dh.ne)
            yield (d, n, "sw", c.dh.sw)
            yield from walk(c.dh.nw, d + 1, 2 * n)
            yield from walk(c.dh.se, d + 1, 2 * n + 1)
        yield from walk(ctypes.pointer(self.root), 0, 0)

    def lr_arrays(self, node_ptr):
        lr = node_ptr.contents.dlr
        h, r = node_ptr.contents.size, lr.rank
        u = np.ctypeslib.as_array(lr.u, shape=(r * h,)).reshape(r, h).T.copy() if r else np.zeros((h, 0))
        v = np.ctypeslib.as_array(lr.v, shape=(r * h,)).reshape(r, h).T.copy() if r else np.zeros((h, 0))
        s = np.ctypeslib.as_array(lr.sig, shape=(r,)).copy() if r else np.zeros(0)
        return u, s, v

    def depth(self):
        d, node = 0, self.root
        while node.type == PLR_HMTX:
            node = node.dh.nw.contents
            d += 1
        return d


def gepmv(trans, A, x, alpha=1.0, beta=0.0, y=None, incx=1, incy=1, api=None):
    """Call dplrgepmv on numpy vectors (x, y may be strided by incx/incy)."""
    f = _fn(api, "dplrgepmv")
    m = A.m
    x = np.ascontiguousarray(x, dtype=np.float64)
    if y is None:
        y = np.zeros(1 + (m - 1) * incy if m else 0)
    else:
        y = np.array(y, dtype=np.float64, copy=True)
    a, b = c_double(alpha), c_double(beta)
    f(c_char(trans.encode()), m, byref(a), A.ref, _ptr(x), incx, byref(b), _ptr(y), incy)
    return y


def gepdmm(transa, transb, A, B, n, alpha=1.0, beta=0.0, C=None, ldb=None, ldc=None, api=None):
    """Call dplrgepdmm; B and C are Fortran-ordered 2-D arrays (leading dims from shapes)."""
    f = _fn(api, "dplrgepdmm")
    m = A.m
    B = np.asfortranarray(B, dtype=np.float64)
    ldb = ldb or max(B.shape[0], 1)
    if C is None:
        C = np.zeros((m, n), order="F")
    else:
        C = np.array(C, dtype=np.float64, order="F", copy=True)
    ldc = ldc or max(C.shape[0], 1)
    a, b = c_double(alpha), c_double(beta)
    f(c_char(transa.encode()), c_char(transb.encode()), m, n, byref(a), A.ref, _ptr(B), ldb,
      byref(b), _ptr(C), ldc)
    return C


def geinv(A, eps=0.0, api=None):
    """B = inv(A) via dplrgeinv; returns (B tree, info)."""
    f = _fn(api, "dplrgeinv")
    B = PlrTree(A.m)
    info = c_int(12345)
    f(A.m, A.ref, B.ref, eps, byref(info))
    return B, info.value


def geprk(A, u, v, s, alpha=1.0, eps=0.0, api=None):
    """A += alpha*u*diag(s)*v^T in place via dplrgeprk."""
    f = _fn(api, "dplrgeprk")
    u = np.asfortranarray(u, dtype=np.float64)
    v = np.asfortranarray(v, dtype=np.float64)
    s = np.ascontiguousarray(s, dtype=np.float64)
    k = u.shape[1]
    a = c_double(alpha)
    f(A.m, k, byref(a), _ptr(u), max(u.shape[0], 1), _ptr(v), max(v.shape[0], 1), _ptr(s), 1,
      A.ref, eps)


def geppmm(transa, transb, A, B, C, alpha=1.0, beta=0.0, eps=0.0, api=None):
    """C := alpha*op(A)*op(B) + beta*C in place via dplrgeppmm (C a PlrTree)."""
    f = _fn(api, "dplrgeppmm")
    a, b = c_double(alpha), c_double(beta)
    f(c_char(transa.encode()), c_char(transb.encode()), A.m, byref(a), A.ref, B.ref, byref(b),
      C.ref, eps)
    return C


def empty(A, api=None):
    """A new tree of A's structure (dplrempty)."""
    B = PlrTree(A.m)
    _fn(api, "dplrempty")(A.ref, B.ref)
    return B


def copy(A, B=None, api=None):
    """B := A via dplrcpy (a new tree unless B is given)."""
    B = B if B is not None else PlrTree(A.m)
    _fn(api, "dplrcpy")(A.m, A.ref, B.ref)
    return B


def nodes(tree):
    """Every node in pre-order (node, nw, ne, se, sw as the H node lists them):
    (type, size, rank or -1, capacity or -1) tuples, for structure checks."""
    out = []

    def walk(p):
        c = p.contents
        if c.type == PLR_LRMTX:
            out.append((c.type, c.size, c.dlr.rank, c.dlr.capacity))
        else:
            out.append((c.type, c.size, -1, -1))
        if c.type == PLR_HMTX:
            for w in ("nw", "ne", "se", "sw"):
                walk(getattr(c.dh, w))
    walk(ctypes.pointer(tree.root))
    return out
