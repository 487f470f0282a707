"""Device-resident PLR plans (include/dplrx.h) driven with torch device tensors.

torch only provides device memory and the stream; every operation is a call
into libdplr_hip.so.
"""
import ctypes
from ctypes import byref, c_int, c_void_p

from ._lib import check, lib


def _stream(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))


def _dptr(t):
    return c_void_p(t.data_ptr())


def _ld(t):
    """Leading dimension of an (n, m) row-major tensor seen as a column-major
    m x n block (torch reports stride 1 for a leading dimension of size 1)."""
    return t.stride(0) if t.shape[0] > 1 else t.shape[1]


class Plan:
    """An uploaded PLR matrix (level-wise FP64 arrays in HBM)."""

    def __init__(self, tree):
        p = c_void_p()
        check(lib.dplrx_plan_create(tree.m, tree.ref, byref(p)), "dplrx_plan_create")
        self.p = p
        N, L, b = c_int(), c_int(), c_int()
        R = (c_int * 32)()
        lib.dplrx_plan_info(self.p, byref(N), byref(L), byref(b), R)
        self.m, self.nlvl, self.leaf = N.value, L.value, b.value
        self.ranks = [R[d] for d in range(self.nlvl)]

    @property
    def nbytes(self):
        return int(lib.dplrx_plan_bytes(self.p))

    def close(self):
        if getattr(self, "p", None):
            lib.dplrx_plan_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def gepmv(self, trans, alpha, x, beta, y, stream=None):
        check(lib.dplrx_gepmv(self.p, trans.encode(), alpha, _dptr(x), beta, _dptr(y),
                              _stream(stream)), "dplrx_gepmv")
        return y

    def gepdmm(self, trans, alpha, B, beta, C, stream=None):
        """C := alpha*op(A)*B + beta*C; B, C are (n, m) row-major torch tensors,
        i.e. column-major m x n blocks with leading dimension m."""
        n, m = B.shape
        check(lib.dplrx_gepdmm(self.p, trans.encode(), n, alpha, _dptr(B), _ld(B), beta,
                               _dptr(C), _ld(C), _stream(stream)), "dplrx_gepdmm")
        return C

    def factor(self, stream=None):
        """FP64 factorization in HBM; returns info (0, or > 0 when singular)."""
        info = c_int(0)
        check(lib.dplrx_factor(self.p, byref(info), _stream(stream)), "dplrx_factor")
        return info.value

    def solve(self, trans, B, X, stream=None):
        """X := op(A)^{-1} B; B, X are (n, m) row-major torch tensors."""
        n, m = B.shape
        check(lib.dplrx_solve(self.p, trans.encode(), n, _dptr(B), _ld(B), _dptr(X),
                              _ld(X), _stream(stream)), "dplrx_solve")
        return X

    # ---- level-wise layout and adjoints -------------------------------------

    def layout(self):
        """Offsets of the level-wise arrays: (col_off, s_off) where level d's
        U/V columns are [col_off[d], col_off[d+1]) of the (cols, m) arrays and
        its sig entries are [s_off[d], s_off[d+1]), segment q at q * R_d."""
        col, so = [0], [0]
        for d, R in enumerate(self.ranks):
            col.append(col[-1] + R)
            so.append(so[-1] + (2 << d) * R)
        return col, so

    def grad_outer(self, alpha, Ga, Gb, stream=None):
        """Gradient of <alpha * Ga Gb^T, A> with respect to the factors, as
        (dU, dV, dS, dD) device tensors in the plan's layouts: dU, dV of shape
        (sum R_d, m) (row k = column k of the level-wise array), dS of
        sum_d 2^(d+1) R_d entries, dD of shape (m/leaf, leaf, leaf) with each
        leaf column-major (dD[l][j][i] = d A(l*leaf + i, l*leaf + j))."""
        import torch
        col, so = self.layout()
        dev = Ga.device
        dU = torch.zeros(col[-1], self.m, dtype=torch.float64, device=dev)
        dV = torch.zeros_like(dU)
        dS = torch.zeros(max(so[-1], 1), dtype=torch.float64, device=dev)
        dD = torch.zeros(self.m // self.leaf, self.leaf, self.leaf, dtype=torch.float64, device=dev)
        n = Ga.shape[0]
        check(lib.dplrx_grad_outer(self.p, n, alpha, _dptr(Ga), _ld(Ga), _dptr(Gb), _ld(Gb),
                                   _dptr(dU), _dptr(dV), _dptr(dS), _dptr(dD), _stream(stream)),
              "dplrx_grad_outer")
        return dU, dV, dS[:so[-1]], dD

    def grad_outer_f32(self, alpha, Ga, Gb, stream=None):
        """grad_outer in FP32 on the plan's FP32 factors (factor_single("f32")):
        Ga, Gb float32 device tensors; returns float32 (dU, dV, dS, dD)."""
        import torch
        col, so = self.layout()
        dev = Ga.device
        f = dict(dtype=torch.float32, device=dev)
        dU = torch.zeros(col[-1], self.m, **f)
        dV = torch.zeros_like(dU)
        dS = torch.zeros(max(so[-1], 1), **f)
        dD = torch.zeros(self.m // self.leaf, self.leaf, self.leaf, **f)
        n = Ga.shape[0]
        check(lib.dplrx_grad_outer_f32(self.p, n, alpha, _dptr(Ga), _ld(Ga), _dptr(Gb), _ld(Gb),
                                       _dptr(dU), _dptr(dV), _dptr(dS), _dptr(dD),
                                       _stream(stream)), "dplrx_grad_outer_f32")
        return dU, dV, dS[:so[-1]], dD

    # ---- single precision ------------------------------------------------------

    SP_TYPES = {"f32": 1, "bf16": 2}

    def factor_single(self, dtype="f32", stream=None):
        """FP32 (or BF16 factor matrices) copies of the factors, after the FP64
        factorization (run first when missing); returns info."""
        info = c_int(0)
        check(lib.dplrx_factor_single(self.p, self.SP_TYPES[dtype], byref(info), _stream(stream)),
              "dplrx_factor_single")
        return info.value

    def gepdmm_single(self, trans, alpha, B, beta, C, dtype="f32", stream=None):
        """C := alpha*op(A)*B + beta*C in single precision; B, C float32 (n, m)."""
        n, m = B.shape
        check(lib.dplrx_gepdmm_single(self.p, self.SP_TYPES[dtype], trans.encode(), n, alpha,
                                      _dptr(B), _ld(B), beta, _dptr(C), _ld(C), _stream(stream)),
              "dplrx_gepdmm_single")
        return C

    def solve_single(self, trans, B, X, dtype="f32", stream=None):
        """X := op(A)^{-1} B in single precision; B, X float32 (n, m)."""
        n, m = B.shape
        check(lib.dplrx_solve_single(self.p, self.SP_TYPES[dtype], trans.encode(), n, _dptr(B),
                                     _ld(B), _dptr(X), _ld(X), _stream(stream)),
              "dplrx_solve_single")
        return X

    LP_TYPES = {"f16": 1, "bf16": 2, "f16x2": 3}

    def factor_mixed(self, dtype="f16", stream=None):
        """Mixed-precision factors (f16, bf16, or f16x2: hi + lo f16 pairs);
        returns info."""
        info = c_int(0)
        check(lib.dplrx_factor_mixed(self.p, self.LP_TYPES[dtype], byref(info), _stream(stream)),
              "dplrx_factor_mixed")
        return info.value

    def solve_refined(self, B, X, iters, with_residual=False, stream=None):
        """X := A^{-1} B with `iters` refined mixed-precision solves.  Returns the
        per-column max|B - A X| when with_residual (synchronises)."""
        import numpy as np
        n, m = B.shape
        res = np.zeros(n) if with_residual else None
        rp = res.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if with_residual else None
        check(lib.dplrx_solve_refined(self.p, n, _dptr(B), _ld(B), _dptr(X), _ld(X),
                                      iters, rp, _stream(stream)), "dplrx_solve_refined")
        return res

    def solve_refined_tol(self, B, X, tol=1e-10, maxit=8, stream=None):
        """X := A^{-1} B refined until the estimated relative error of every
        column is <= tol (at most maxit solves).  Returns (solves made,
        error estimate); synchronises."""
        n, m = B.shape
        it, err = c_int(0), ctypes.c_double(0.0)
        check(lib.dplrx_solve_refined_tol(self.p, n, _dptr(B), _ld(B), _dptr(X), _ld(X), tol,
                                          maxit, byref(it), byref(err), _stream(stream)),
              "dplrx_solve_refined_tol")
        return it.value, err.value


    def apply_solve_refined_tol(self, Xa, Ya, B, X, tol=1e-10, maxit=8, stream=None):
        """One mixed-precision step: Ya := A Xa (FP64) and X := A^{-1} B
        refined to tol, the apply sharing the first residual product.
        Returns (solves made, error estimate); synchronises."""
        n, m = B.shape
        if Xa.shape != B.shape or Ya.shape != B.shape:
            raise ValueError("Xa, Ya and B must have the same shape")
        it, err = c_int(0), ctypes.c_double(0.0)
        check(lib.dplrx_apply_solve_refined_tol(self.p, n, _dptr(Xa), _ld(Xa), _dptr(Ya), _ld(Ya),
                                                _dptr(B), _ld(B), _dptr(X), _ld(X), tol, maxit,
                                                byref(it), byref(err), _stream(stream)),
              "dplrx_apply_solve_refined_tol")
        return it.value, err.value


class Batch:
    """Independent plans applied or solved together: one kernel launch per
    phase covers every matrix, replayed from a HIP graph (dplrx_batch_*).
    The plans must stay alive as long as the batch."""

    def __init__(self, plans):
        self.plans = list(plans)
        arr = (c_void_p * len(self.plans))(*[p.p for p in self.plans])
        b = c_void_p()
        check(lib.dplrx_batch_create(len(self.plans), arr, byref(b)), "dplrx_batch_create")
        self.b = b

    def close(self):
        if getattr(self, "b", None):
            lib.dplrx_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _io(self, X, Y):
        n = len(self.plans)
        xs = (c_void_p * n)(*[x.data_ptr() for x in X])
        ys = (c_void_p * n)(*[y.data_ptr() for y in Y])
        lx = (c_int * n)(*[_ld(x) for x in X])
        ly = (c_int * n)(*[_ld(y) for y in Y])
        return xs, lx, ys, ly

    PREC = {"f64": 0, "f32": 1, "bf16": 2}

    def gepdmm(self, X, Y, prec="f64", trans="N", alpha=1.0, beta=0.0, stream=None):
        """Y[i] := alpha op(A_i) X[i] + beta Y[i]; X[i], Y[i] are (nrhs, N_i)
        device tensors (float64 for "f64", float32 for "f32" and "bf16")."""
        nrhs = X[0].shape[0]
        xs, lx, ys, ly = self._io(X, Y)
        if prec == "f64" and trans == "N" and alpha == 1.0 and beta == 0.0:
            check(lib.dplrx_batch_gepdmm(self.b, nrhs, xs, lx, ys, ly, _stream(stream)),
                  "dplrx_batch_gepdmm")
        else:
            check(lib.dplrx_batch_op(self.b, self.PREC[prec], 0, trans.encode(), nrhs, xs, lx, ys,
                                     ly, alpha, beta, _stream(stream)), "dplrx_batch_op")
        return Y

    def solve(self, B, X, prec="f64", trans="N", stream=None):
        """X[i] := op(A_i)^{-1} B[i] (every plan factored in that precision)."""
        nrhs = B[0].shape[0]
        bs, lb, xs, lx = self._io(B, X)
        if prec == "f64" and trans == "N":
            check(lib.dplrx_batch_solve(self.b, nrhs, bs, lb, xs, lx, _stream(stream)),
                  "dplrx_batch_solve")
        else:
            check(lib.dplrx_batch_op(self.b, self.PREC[prec], 1, trans.encode(), nrhs, bs, lb, xs,
                                     lx, 0.0, 0.0, _stream(stream)), "dplrx_batch_op")
        return X

    def solve_refined(self, B, X, iters, stream=None):
        """X[i] := A_i^{-1} B[i] by `iters` BF16 solves refined with FP32
        residuals (plans prepared with factor_single "f32" and "bf16")."""
        nrhs = B[0].shape[0]
        bs, lb, xs, lx = self._io(B, X)
        check(lib.dplrx_batch_solve_refined(self.b, nrhs, bs, lb, xs, lx, iters, _stream(stream)),
              "dplrx_batch_solve_refined")
        return X

    TILES = (128, 256, 512, 1024, 2048, 4096)

    def set_tile(self, tile):
        """Rows per reduction task (one of TILES), without timing."""
        check(lib.dplrx_batch_set_tile(self.b, tile), "dplrx_batch_set_tile")

    def tune(self, X, Y, reps=3, prec="f64", stream=None):
        """Autotuner: times apply + solve of this batch (X read, Y written) in
        precision prec for each candidate tile shape (rows per reduction task)
        and keeps the fastest.  Returns (tile, {tile: ms per repetition})."""
        import numpy as np
        nrhs = X[0].shape[0]
        xs, lx, ys, ly = self._io(X, Y)
        tile = c_int(0)
        ms = np.zeros(len(self.TILES))
        check(lib.dplrx_batch_tune_prec(self.b, self.PREC[prec], nrhs, xs, lx, ys, ly, reps,
                                        byref(tile),
                                        ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                        _stream(stream)), "dplrx_batch_tune_prec")
        return tile.value, dict(zip(self.TILES, ms.tolist()))
