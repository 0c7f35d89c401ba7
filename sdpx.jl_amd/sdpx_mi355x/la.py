"""Host mirror of the LA provider seam's host-buffer operations (la_backend.jl:1529-1769).

``MI355XLABackend`` carries the la_* methods SDPX's dense route may call on a
provider backend -- la_chol!, la_trsm!, la_trsv_lower!, la_trsv_transpose!,
la_syrk!, la_mul_owned!, la_axpby_owned!, la_dot, la_norminf -- with the
reference's in-place conventions, each computed on the device through the
C-ABI (``sdpx_la_*``).  Arrays are float64 numpy arrays in Fortran order (the
reference's column-major layout); in-place operands must be writable.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dptr


def _fortran(a, name, writable=False):
    """A float64 vector, or a column-major matrix (unit row stride; a leading dimension of at least
    its rows, as a view of the top rows of a larger Fortran array has)."""
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64):
        raise TypeError(f"{name}: a float64 numpy array is required")
    if writable and not a.flags.writeable:
        raise ValueError(f"{name}: the array must be writable")
    if a.ndim == 1:
        if a.strides[0] != 8:
            raise ValueError(f"{name}: a contiguous vector is required")
        return a
    if a.ndim != 2 or (a.shape[0] > 1 and a.strides[0] != 8) or a.strides[1] % 8 or \
            (a.shape[1] > 1 and a.strides[1] // 8 < a.shape[0]):
        raise ValueError(f"{name}: a column-major (Fortran-order) matrix is required")
    return a


def _ld(a):
    if a.ndim == 2 and a.shape[1] > 1:
        return max(a.strides[1] // 8, a.shape[0], 1)
    return max(a.shape[0], 1)


class MI355XLABackend:
    """One device handle; the methods follow the reference's la_* signatures (backend argument dropped)."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = C.c_void_p()
        check(self.lib.sdpx_la_create(device, C.byref(h)), "sdpx_la_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.sdpx_la_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def la_chol(self, A) -> bool:
        """la_chol! (la_backend.jl:1612-1622): lower Cholesky in place; False when not positive definite."""
        A = _fortran(A, "la_chol", True)
        n = A.shape[0]
        if A.shape != (n, n):
            raise ValueError("la_chol: A must be square")
        return check(self.lib.sdpx_la_chol(self.h, n, dptr(A), _ld(A)), "sdpx_la_chol") == 0

    def la_trsm(self, L, X):
        """la_trsm! (:1634): X <- L^-1 X (X a matrix or a vector)."""
        L = _fortran(L, "la_trsm L")
        X = _fortran(X, "la_trsm X", True)
        n = L.shape[0]
        nrhs = 1 if X.ndim == 1 else X.shape[1]
        if L.shape != (n, n) or X.shape[0] != n:
            raise ValueError("la_trsm: L n x n and X with n rows required")
        check(self.lib.sdpx_la_trsm(self.h, n, nrhs, dptr(L), _ld(L), dptr(X), max(n, 1)), "sdpx_la_trsm")
        return X

    def _trsv(self, L, x, trans):
        L = _fortran(L, "la_trsv L")
        n = L.shape[0]
        if L.shape != (n, n) or x.shape != (n,) or not x.flags.writeable:
            raise ValueError("la_trsv: L n x n and a writable x of n entries required")
        check(self.lib.sdpx_la_trsv(self.h, n, dptr(L), _ld(L), dptr(x), trans), "sdpx_la_trsv")
        return x

    def la_trsv_lower(self, L, x):
        """la_trsv_lower! (:1642): x <- L^-1 x."""
        return self._trsv(L, x, 0)

    def la_trsv_transpose(self, L, x):
        """la_trsv_transpose! (:1648-1653): x <- L^-T x."""
        return self._trsv(L, x, 1)

    def la_syrk(self, S, P, alpha, beta):
        """la_syrk! (:1595-1604): lower triangle of S <- alpha P'P + beta S (BLAS syrk 'L', 'T')."""
        S = _fortran(S, "la_syrk S", True)
        P = _fortran(P, "la_syrk P")
        k, n = P.shape
        if S.shape != (n, n):
            raise ValueError("la_syrk: S must be n x n with P k x n")
        check(self.lib.sdpx_la_syrk(self.h, n, k, float(alpha), dptr(P), _ld(P), float(beta), dptr(S), _ld(S)),
              "sdpx_la_syrk")
        return S

    def la_mul_owned(self, Cm, A, B, alpha=1.0, beta=0.0, ta=False, tb=False):
        """la_mul_owned! (:1556-1577): C <- alpha op(A) op(B) + beta C, op = transpose when ta / tb."""
        Cm = _fortran(Cm, "la_mul C", True)
        A = _fortran(A, "la_mul A")
        B = _fortran(B, "la_mul B")
        m, n = Cm.shape
        k = A.shape[0] if ta else A.shape[1]
        if (A.shape[1] if ta else A.shape[0]) != m or (B.shape[1] if tb else B.shape[0]) != k or \
                (B.shape[0] if tb else B.shape[1]) != n:
            raise ValueError("la_mul: operand shapes do not match")
        check(self.lib.sdpx_la_gemm(self.h, int(ta), int(tb), m, n, k, float(alpha), dptr(A), _ld(A), dptr(B), _ld(B),
                                    float(beta), dptr(Cm), _ld(Cm)), "sdpx_la_gemm")
        return Cm

    def la_axpby_owned(self, alpha, X, beta, Y):
        """la_axpby_owned! (:1659-1686): Y <- alpha X + beta Y."""
        if X.shape != Y.shape or not Y.flags.writeable:
            raise ValueError("la_axpby: X and a writable Y of the same shape required")
        if not (Y.flags.c_contiguous or Y.flags.f_contiguous) or Y.dtype != np.float64:
            raise ValueError("la_axpby: Y must be a contiguous float64 array")
        order = "C" if Y.flags.c_contiguous else "F"
        x = np.asarray(X, dtype=np.float64).ravel(order=order)
        y = Y.reshape(-1, order=order)
        check(self.lib.sdpx_la_axpby(self.h, y.size, float(alpha), dptr(np.ascontiguousarray(x)), float(beta),
                                     dptr(y)), "sdpx_la_axpby")
        return Y

    def la_dot(self, x, y) -> float:
        """la_dot (:1529)."""
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        y = np.ascontiguousarray(y, dtype=np.float64).ravel()
        if x.size != y.size:
            raise ValueError("la_dot: lengths differ")
        out = C.c_double(0.0)
        check(self.lib.sdpx_la_dot(self.h, x.size, dptr(x), dptr(y), C.byref(out)), "sdpx_la_dot")
        return out.value

    def la_norminf(self, x) -> float:
        """la_norminf (:1537)."""
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        out = C.c_double(0.0)
        check(self.lib.sdpx_la_norminf(self.h, x.size, dptr(x), C.byref(out)), "sdpx_la_norminf")
        return out.value
