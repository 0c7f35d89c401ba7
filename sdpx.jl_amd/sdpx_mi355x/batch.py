"""Batched solver: many independent single-block SDPs solved in lock step on one GPU.

Host binding of the sdpx_batch_* entry points (include/sdpx_mi355x.h).  Each
member is a core-form SDP  min c'x  s.t.  sum_i x_i A_i - C >= 0  with one PSD
block of side k <= 64 and m <= 2048 constraints; every Newton-step phase is
one launch over the members still iterating, the per-member interior-point
control runs in the library.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, dptr
from .policy import SolverOptions

STATUS = {0: "Optimal", 1: "IterLimit", 2: "NumericalBreakdown", 3: "Stalled", 4: "InsufficientPrecision",
          5: "MaxRestartsExceeded"}


class BatchOptions(C.Structure):
    """sdpx_batch_options_t."""

    _fields_ = [
        ("beta", C.c_double), ("gamma", C.c_double), ("eps_gap", C.c_double), ("eps_primal", C.c_double),
        ("eps_dual", C.c_double), ("min_step", C.c_double), ("omega_step", C.c_double),
        ("stall_tolerance", C.c_double), ("refine_tol", C.c_double), ("adaptive_sigma_max", C.c_double),
        ("iter_max", C.c_int), ("max_restarts", C.c_int), ("max_centering", C.c_int),
        ("stall_iterations", C.c_int), ("restart", C.c_int), ("refine_policy", C.c_int),
        ("refine_steps", C.c_int), ("refine_max_steps", C.c_int), ("adaptive", C.c_int), ("step_rule", C.c_int),
        ("precision", C.c_int), ("cond_limit", C.c_double), ("refine_target", C.c_double),
        ("adjoint_mu", C.c_double), ("mixed_max_refine", C.c_int),
    ]

    @classmethod
    def from_solver_options(cls, o: SolverOptions, precision: int | None = None) -> "BatchOptions":
        """precision None follows o.precision ("fp64" -> 0, "mixed" -> 1); the mixed route takes
        o.mixed_cond_limit, o.mixed_refine_target and o.mixed_batch_max_refine."""
        b = cls()
        check(_lib.load().sdpx_batch_default_options(C.addressof(b), o.eps_gap), "sdpx_batch_default_options")
        for name in ("beta", "gamma", "eps_gap", "eps_primal", "eps_dual", "min_step", "omega_step",
                     "stall_tolerance", "refine_tol", "adaptive_sigma_max", "iter_max", "max_restarts",
                     "max_centering", "stall_iterations", "refine_steps", "refine_max_steps"):
            setattr(b, name, getattr(o, name))
        b.restart = int(o.restart)
        b.refine_policy = 1 if o.refine_policy == "fixed" else 0
        b.adaptive = 1 if o.parameter_strategy == "adaptive" else 0
        b.step_rule = {"auto": 0, "backtrack": 1, "fraction_to_boundary": 2}[o.step_rule]
        if precision is None:
            if o.precision == "mixed_bf16x3":
                raise ValueError("the batched route assembles in FP64 or FP32 (precision 'fp64' or 'mixed')")
            precision = 1 if o.precision == "mixed" else 0
        b.precision = precision
        b.cond_limit = o.mixed_cond_limit
        b.refine_target = o.mixed_refine_target
        b.mixed_max_refine = o.mixed_batch_max_refine
        return b


@dataclass
class BatchResult:
    status: list
    iterations: np.ndarray
    pobj: np.ndarray
    dobj: np.ndarray
    x: list
    X: list
    Y: list
    launched: int

    @property
    def gap_rel(self) -> np.ndarray:
        return np.abs(self.pobj - self.dobj) / np.maximum(1.0, (np.abs(self.pobj) + np.abs(self.dobj)) / 2)


def _members(name, arrays, sizes, B):
    """Per-member arrays flattened after checking each against its expected element count."""
    arrays = list(arrays)
    if len(arrays) != B:
        raise ValueError(f"{name}: {len(arrays)} members given, the batch has {B}")
    out = []
    for p, (a, n) in enumerate(zip(arrays, sizes)):
        a = np.asarray(a, dtype=np.float64)
        if a.size != int(n):
            raise ValueError(f"{name}: member {p} has {a.size} values, expected {int(n)}")
        out.append(a)
    return out


class BatchSolver:
    """Device-resident batch of single-block SDPs (block sides ``k``, constraint counts ``m``)."""

    def __init__(self, k, m, device: int = 0):
        lib = _lib.load()
        self.k = np.ascontiguousarray(k, dtype=np.int32)
        self.m = np.ascontiguousarray(m, dtype=np.int32)
        if self.k.shape != self.m.shape or self.k.ndim != 1:
            raise ValueError("k and m must be 1-D of equal length")
        self.B = int(self.k.size)
        h = C.c_void_p()
        check(lib.sdpx_batch_create(device, self.B, self.k.ctypes.data_as(_lib._ip),
                                    self.m.ctypes.data_as(_lib._ip), C.byref(h)), "sdpx_batch_create")
        self._h = h
        self._lib = lib

    def close(self):
        if self._h:
            self._lib.sdpx_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def memory_bytes(self) -> int:
        v = C.c_longlong(0)
        check(self._lib.sdpx_batch_memory(self._h, C.byref(v)), "sdpx_batch_memory")
        return v.value

    # ---- data -------------------------------------------------------------------------------
    def set_dense(self, Av_list, C_list, c_list) -> None:
        """Av_list[p]: (m_p, k_p^2) rows = vec(A_i) (column-major), C_list[p]: (k_p, k_p), c_list[p]: (m_p,)."""
        k2 = self.k.astype(np.int64) ** 2
        Av_list = _members("set_dense Av", Av_list, k2 * self.m, self.B)
        C_list = _members("set_dense C", C_list, k2, self.B)
        c_list = _members("set_dense c", c_list, self.m, self.B)
        Av = np.concatenate([np.ascontiguousarray(a).ravel() for a in Av_list])
        Cc = np.concatenate([c_.ravel(order="F") for c_ in C_list])
        cc = np.concatenate([v.ravel() for v in c_list])
        check(self._lib.sdpx_batch_set_dense(self._h, dptr(Av), dptr(Cc), dptr(cc)), "sdpx_batch_set_dense")

    def set_coo(self, rows, cols, vals, C_list, c_list) -> None:
        """rows/cols/vals: per constraint of every member (concatenated), upper-triangle COO."""
        k2 = self.k.astype(np.int64) ** 2
        C_list = _members("set_coo C", C_list, k2, self.B)
        c_list = _members("set_coo c", c_list, self.m, self.B)
        if not (len(rows) == len(cols) == len(vals) == int(self.m.sum())):
            raise ValueError(f"set_coo: rows, cols and vals need one entry list per constraint ({int(self.m.sum())})")
        for i, (r_, c_, v_) in enumerate(zip(rows, cols, vals)):
            if not (len(r_) == len(c_) == len(v_)):
                raise ValueError(f"set_coo: constraint {i} has rows, cols and vals of different lengths")
        counts = np.array([len(r) for r in rows], dtype=np.int64)
        ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        r = np.concatenate(rows).astype(np.int32) if rows else np.zeros(0, np.int32)
        c = np.concatenate(cols).astype(np.int32) if cols else np.zeros(0, np.int32)
        v = np.concatenate(vals).astype(np.float64) if vals else np.zeros(0)
        Cc = np.concatenate([np.asarray(c_, dtype=np.float64).ravel(order="F") for c_ in C_list])
        cc = np.concatenate([np.asarray(x, dtype=np.float64).ravel() for x in c_list])
        check(self._lib.sdpx_batch_set_coo(self._h, ptr.ctypes.data_as(C.POINTER(C.c_longlong)),
                                           r.ctypes.data_as(_lib._ip), c.ctypes.data_as(_lib._ip), dptr(v),
                                           dptr(Cc), dptr(cc)), "sdpx_batch_set_coo")

    def set_block_coo(self, coo_list, C_list, c_list) -> None:
        """coo_list[p] = (ptr, lin, row, col, val): member p's SparseBlockCOO as the reference stores it
        (types/constraints.jl:59-109: 1-based Int32, both triangles, lin = (col-1) k + row)."""
        k2 = self.k.astype(np.int64) ** 2
        C_list = _members("set_block_coo C", C_list, k2, self.B)
        c_list = _members("set_block_coo c", c_list, self.m, self.B)
        coo_list = list(coo_list)
        if len(coo_list) != self.B:
            raise ValueError(f"set_block_coo: {len(coo_list)} members given, the batch has {self.B}")
        parts = [[], [], [], [], []]
        for p, (ptr, lin, row, col, val) in enumerate(coo_list):
            ptr = np.asarray(ptr, dtype=np.int32)
            n = int(ptr[-1]) - 1 if ptr.size else -1
            if ptr.size != int(self.m[p]) + 1 or not (len(lin) == len(row) == len(col) == len(val) == n):
                raise ValueError(f"set_block_coo: member {p} needs m+1 positions and ptr[end]-1 entries")
            for dst, a in zip(parts, (ptr, lin, row, col)):
                dst.append(np.asarray(a, dtype=np.int32))
            parts[4].append(np.asarray(val, dtype=np.float64))
        ptr, lin, row, col = (np.concatenate(x) for x in parts[:4])
        val = np.concatenate(parts[4])
        Cc = np.concatenate([c_.ravel(order="F") for c_ in C_list])
        cc = np.concatenate([v.ravel() for v in c_list])
        ip = _lib._ip
        check(self._lib.sdpx_batch_set_block_coo(self._h, ptr.ctypes.data_as(ip), lin.ctypes.data_as(ip),
                                                 row.ctypes.data_as(ip), col.ctypes.data_as(ip), dptr(val),
                                                 dptr(Cc), dptr(cc)), "sdpx_batch_set_block_coo")

    def set_start(self, x_list, X_list, Y_list) -> None:
        k2 = self.k.astype(np.int64) ** 2
        x_list = _members("set_start x", x_list, self.m, self.B)
        X_list = _members("set_start X", X_list, k2, self.B)
        Y_list = _members("set_start Y", Y_list, k2, self.B)
        x = np.concatenate([np.asarray(v, dtype=np.float64).ravel() for v in x_list])
        X = np.concatenate([np.asarray(v, dtype=np.float64).ravel(order="F") for v in X_list])
        Y = np.concatenate([np.asarray(v, dtype=np.float64).ravel(order="F") for v in Y_list])
        check(self._lib.sdpx_batch_set_start(self._h, dptr(x), dptr(X), dptr(Y)), "sdpx_batch_set_start")

    # ---- solve ----------------------------------------------------------------------------------
    def solve(self, o: SolverOptions | BatchOptions | None = None) -> BatchResult:
        bo = o if isinstance(o, BatchOptions) else BatchOptions.from_solver_options(o or SolverOptions())
        it = C.c_int(0)
        check(self._lib.sdpx_batch_solve(self._h, C.addressof(bo), C.byref(it)), "sdpx_batch_solve")
        return self.result(it.value)

    def result(self, launched: int = 0) -> BatchResult:
        B = self.B
        status = np.empty(B, dtype=np.int32)
        iters = np.empty(B, dtype=np.int32)
        pobj, dobj = np.empty(B), np.empty(B)
        x = np.empty(int(self.m.sum()))
        k2 = (self.k.astype(np.int64) ** 2)
        X, Y = np.empty(int(k2.sum())), np.empty(int(k2.sum()))
        check(self._lib.sdpx_batch_get_result(self._h, status.ctypes.data_as(_lib._ip), iters.ctypes.data_as(_lib._ip),
                                              dptr(pobj), dptr(dobj), dptr(x), dptr(X), dptr(Y)),
              "sdpx_batch_get_result")
        xs = np.split(x, np.cumsum(self.m)[:-1])
        Xs = [b.reshape(int(kk), int(kk), order="F") for b, kk in zip(np.split(X, np.cumsum(k2)[:-1]), self.k)]
        Ys = [b.reshape(int(kk), int(kk), order="F") for b, kk in zip(np.split(Y, np.cumsum(k2)[:-1]), self.k)]
        return BatchResult([STATUS[int(s)] for s in status], iters, pobj, dobj, xs, Xs, Ys, launched)

    # ---- differentiation ---------------------------------------------------------------------
    def restore_checkpoint(self):
        """Replace each member's iterate by its differentiation checkpoint -> (restored count, mu per member)."""
        n = C.c_int(0)
        mu = np.empty(self.B)
        check(self._lib.sdpx_batch_restore_checkpoint(self._h, C.byref(n), dptr(mu)), "sdpx_batch_restore_checkpoint")
        return n.value, mu

    def center(self, iterations: int = 4) -> None:
        """Newton steps on XY = mu I (mu fixed at entry) from the current iterate."""
        check(self._lib.sdpx_batch_center(self._h, int(iterations)), "sdpx_batch_center")

    GRAM_ORDERS = ("k_descending", "member_major", "xcd_bins")

    def autotune(self, repeats: int = 3) -> dict:
        """Time the grouped Schur Gram launch at the start point under each tile order and keep the
        fastest (sdpx_batch_autotune); needs set_start first."""
        out = np.zeros(4)
        check(self._lib.sdpx_batch_autotune(self._h, int(repeats), dptr(out)), "sdpx_batch_autotune")
        return {"gram_order": self.GRAM_ORDERS[int(out[0])],
                "gram_ms": {name: round(float(v), 4) for name, v in zip(self.GRAM_ORDERS, out[1:])}}

    def set_gram_order(self, order: str) -> None:
        check(self._lib.sdpx_batch_set_gram_order(self._h, self.GRAM_ORDERS.index(order)), "sdpx_batch_set_gram_order")

    def center_at(self, mu, iterations: int = 4) -> None:
        """Newton steps on XY = mu[p] I per member (NaN: the member's current <X,Y>/k)."""
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        if mu.shape != (self.B,):
            raise ValueError(f"center_at: mu needs {self.B} entries")
        check(self._lib.sdpx_batch_center_at(self._h, int(iterations), dptr(mu)), "sdpx_batch_center_at")

    def adjoint_limit(self, G_list, h_list, ratio: float = 0.1, iterations: int = 6):
        """The optimum's derivative as the mu -> 0 limit of the central-path adjoint: from each member's
        checkpoint (mu_ck), the adjoint D1 at the central-path point of mu_ck and D2 at ratio * mu_ck,
        combined by one Richardson step for a derivative linear in mu, (D2 - ratio D1) / (1 - ratio).
        Leaves the iterate at the second point.  Returns (dC list, db list, mu_ck)."""
        if not 0.0 < ratio < 1.0:
            raise ValueError("adjoint_limit: ratio must lie in (0, 1)")
        _, mu = self.restore_checkpoint()
        self.center(iterations)
        dC1, db1 = self.adjoint(G_list, h_list)
        self.center_at(np.where(np.isnan(mu), np.nan, ratio * mu), iterations)
        dC2, db2 = self.adjoint(G_list, h_list)
        w = 1.0 / (1.0 - ratio)
        dC = [(b - ratio * a) * w for a, b in zip(dC1, dC2)]
        db = [(b - ratio * a) * w for a, b in zip(db1, db2)]
        return dC, db, mu

    def adjoint(self, G_list, h_list):
        """dL/dC and dL/db of L = sum_p <G_p, X*_p> + h_p'y*_p (user's dual-form model) at the current iterate.

        Returns (dC list of symmetric k_p x k_p, db list of m_p vectors); members that could not be
        factored get NaN."""
        k2 = self.k.astype(np.int64) ** 2
        G_list = _members("adjoint G", G_list, k2, self.B)
        h_list = _members("adjoint h", h_list, self.m, self.B)
        G = np.concatenate([g.ravel(order="F") for g in G_list])
        hv = np.concatenate([v.ravel() for v in h_list])
        dC = np.empty(int(k2.sum()))
        db = np.empty(int(self.m.sum()))
        check(self._lib.sdpx_batch_adjoint(self._h, dptr(G), dptr(hv), dptr(dC), dptr(db)), "sdpx_batch_adjoint")
        dCs = [b.reshape(int(kk), int(kk), order="F") for b, kk in zip(np.split(dC, np.cumsum(k2)[:-1]), self.k)]
        return dCs, np.split(db, np.cumsum(self.m)[:-1])

    TRACE_COLUMNS = ("iteration", "pobj", "dobj", "p_res", "d_res", "tX", "tY", "sigma", "affine_primal",
                     "affine_dual", "reg_attempts", "refine_steps", "outcome")

    def trace(self, member: int) -> list:
        """Per-step rows of the member's last solve as dicts (outcome: update / centering / restart / end)."""
        n = C.c_int(0)
        check(self._lib.sdpx_batch_get_trace(self._h, member, None, 0, C.byref(n)), "sdpx_batch_get_trace")
        cols = len(self.TRACE_COLUMNS)
        buf = np.empty(max(n.value, 1) * cols)
        check(self._lib.sdpx_batch_get_trace(self._h, member, dptr(buf), n.value, C.byref(n)),
              "sdpx_batch_get_trace")
        rows = buf[: n.value * cols].reshape(n.value, cols)
        names = {0: "update", 1: "centering", 2: "restart", 3: "end"}
        out = []
        for r in rows:
            d = dict(zip(self.TRACE_COLUMNS, r.tolist()))
            d["outcome"] = names[int(d["outcome"])]
            out.append(d)
        return out

    def set_precision(self, precision: int) -> None:
        """Assembly precision of schur_at_start and autotune: 0 FP64, 1 FP32 (mixed precision)."""
        check(self._lib.sdpx_batch_set_precision(self._h, int(precision)), "sdpx_batch_set_precision")

    def mixed_iterations(self) -> np.ndarray:
        """Per member: Newton steps of the last solve taken with the reduced-precision Schur matrix."""
        out = np.zeros(self.B, dtype=np.int32)
        check(self._lib.sdpx_batch_get_mixed_iterations(self._h, out.ctypes.data_as(_lib._ip)),
              "sdpx_batch_get_mixed_iterations")
        return out

    REASONS = {0: "none", 1: "optimal", 2: "block_factor_failed", 3: "schur_not_positive_definite",
               4: "direction_residual", 5: "collapse_near_solution", 6: "collapse_insufficient_progress",
               7: "collapse_exhausted", 8: "nonfinite_iterate", 9: "iteration_limit", 11: "stagnation_no_progress",
               12: "stagnation_too_slow", 13: "stagnation_precision_floor"}

    def reasons(self) -> list:
        """Per member: why the last solve stopped (sdpx_batch_get_reasons)."""
        out = np.zeros(self.B, dtype=np.int32)
        check(self._lib.sdpx_batch_get_reasons(self._h, out.ctypes.data_as(_lib._ip)), "sdpx_batch_get_reasons")
        return [self.REASONS.get(int(v), str(int(v))) for v in out]

    def schur_at_start(self) -> int:
        return check(self._lib.sdpx_batch_schur_at_start(self._h), "sdpx_batch_schur_at_start")

    def schur(self, member: int) -> np.ndarray:
        mm = int(self.m[member])
        S = np.empty(mm * mm)
        check(self._lib.sdpx_batch_get_schur(self._h, member, dptr(S)), "sdpx_batch_get_schur")
        return S.reshape(mm, mm, order="F")
