"""The MI355X dense KKT/Schur backend: one device handle per SDP workspace.

``MI355XDenseBackend`` is the host mirror of the provider a Julia extension
registers behind SDPX's seams (kkt_backend.jl:24-59 ``KKTBackend`` subtypes,
la_backend.jl:508-559 ``instantiate_la_backend``).  Its methods map one-to-one
onto the C-ABI (include/sdpx_mi355x.h); the problem data, iterate, Schur
matrix and factor stay resident in HBM between calls.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import check, dptr


@dataclass
class SDPProblem:
    """Core geometric SDP (types/problems.jl:1-12) with dense panels, n_eq = 0.

    ``Av[l]`` is a C-contiguous (m, k_l^2) array whose row i is vec(A_i^(l)),
    i.e. the bytes of the reference's k^2 x m column-major ``DenseCons.Av[l]``.
    """

    c: np.ndarray
    C: list
    Av: list
    k: list
    # blocks on the sparse Schur route: l -> (ids, SparseBlockCOO); Av[l] is None for them
    sparse: dict = field(default_factory=dict)

    @property
    def m(self) -> int:
        return int(self.c.shape[0])

    @property
    def L(self) -> int:
        return len(self.k)


@dataclass
class SparseBlockCOO:
    """SparseBlockCOO (types/constraints.jl:59-65) as the reference stores it: 1-based Int32
    ``ptr`` over positions, ``lin = (col-1) k + row``, ``row``, ``col`` and ``val``, every
    stored entry of each symmetric coefficient (both triangles)."""

    ptr: np.ndarray
    lin: np.ndarray
    row: np.ndarray
    col: np.ndarray
    val: np.ndarray

    @classmethod
    def build(cls, Al: np.ndarray, order) -> "SparseBlockCOO":
        """build_block_coo (types/constraints.jl:79-109): ``Al`` (m, k, k), positions in
        ``order``, each coefficient's nonzeros in column-major order."""
        Al = np.asarray(Al, dtype=np.float64)
        k = Al.shape[1]
        order = np.asarray(order, dtype=np.int64)
        sub = Al[order].transpose(0, 2, 1).reshape(len(order), k * k)  # row p = vec(A_order[p])
        pos, lin0 = np.nonzero(sub)                                       # position-major, lin ascending
        ptr = np.searchsorted(pos, np.arange(len(order) + 1)) + 1
        i32 = lambda v: np.ascontiguousarray(v, dtype=np.int32)  # noqa: E731
        return cls(i32(ptr), i32(lin0 + 1), i32(lin0 % k + 1), i32(lin0 // k + 1), sub[pos, lin0].copy())


def block_active(Al: np.ndarray) -> np.ndarray:
    """Variables with a structurally nonzero coefficient on the block, ascending (the
    reference's ``active`` / ``schur_order``, test/sparse.jl:206-216)."""
    Al = np.asarray(Al)
    return np.flatnonzero(np.any(Al.reshape(Al.shape[0], -1) != 0.0, axis=1))


def arrow_owner(prob: SDPProblem, max_local: int = 128):
    """ArrowWorkspace's partition (workspace.jl:318-400) as the C-ABI takes it: owner[i] = the one
    block whose active set holds variable i, -1 for a variable on several blocks.  None when the
    block-arrow route does not apply: a variable on no block, every variable global, or a block
    with more than ``max_local`` local variables (the device route's limit)."""
    m = prob.m
    freq = np.zeros(m, dtype=np.int64)
    owner = np.full(m, -1, dtype=np.int32)
    for l in range(prob.L):
        if l in prob.sparse:
            act = np.asarray(prob.sparse[l][0], dtype=np.int64)
        else:
            act = np.flatnonzero(np.any(np.asarray(prob.Av[l]) != 0.0, axis=1))
        freq[act] += 1
        owner[act] = l
    if not np.all(freq > 0) or not np.any(freq == 1):
        return None
    owner[freq > 1] = -1
    counts = np.bincount(owner[owner >= 0], minlength=prob.L)
    if counts.max(initial=0) > max_local:
        return None
    return owner


def ingest(c, A, C, sparse=()) -> SDPProblem:
    """Dense ingest (ingest.jl:680-694): ``A[l]`` is (m, k, k) with A[l][i] = A_i^(l).

    Blocks listed in ``sparse`` are packed as SparseBlockCOO over their active variables
    (the ingest's sparse route, types/constraints.jl:79-109) instead of a dense panel."""
    c = np.ascontiguousarray(c, dtype=np.float64)
    Av, Cs, ks, sp = [], [], [], {}
    for l, (Al, Cl) in enumerate(zip(A, C)):
        Al = np.asarray(Al, dtype=np.float64)
        m, k, k2 = Al.shape
        if k != k2 or m != c.shape[0]:
            raise ValueError("A blocks must be (m, k, k) with m = len(c)")
        if l in sparse:
            ids = block_active(Al)
            if ids.size == 0:
                raise ValueError(f"ingest: sparse block {l} has no active variable")
            sp[l] = (ids, SparseBlockCOO.build(Al, ids))
            Av.append(None)
        else:
            Av.append(np.ascontiguousarray(Al.transpose(0, 2, 1).reshape(m, k * k)))
        Cs.append(np.asfortranarray(np.asarray(Cl, dtype=np.float64)))
        ks.append(int(k))
    return SDPProblem(c=c, C=Cs, Av=Av, k=ks, sparse=sp)


COLD_START_REASONS = {0: "none", 1: "identity_kkt_factorization", 2: "kkt_cold_start_solve",
                      3: "kkt_cold_start_residuals", 4: "psd_shift", 5: "identity_mass", 6: "centering"}


@dataclass
class Equilibration:
    """Ruiz row scales per block, per-variable scales and the objective scale (ingest.jl:1120-1139)."""

    E: list
    s: np.ndarray
    objective_scale: float
    passes: list

    def unequilibrate(self, x, X, Y, y=None):
        """unequilibrate (ingest.jl:1523-1560): the original problem's (x, X, y, Y)."""
        x = np.asarray(x) / self.s
        Xo = [Xl / np.outer(E, E) for Xl, E in zip(X, self.E)]
        Yo = [Yl * np.outer(E, E) for Yl, E in zip(Y, self.E)]
        if self.objective_scale != 1.0:
            Yo = [Yl * self.objective_scale for Yl in Yo]
            y = None if y is None else np.asarray(y) * self.objective_scale
        return x, Xo, y, Yo


def _fortran_concat(blocks) -> np.ndarray:
    return np.concatenate([np.asarray(b, dtype=np.float64).ravel(order="F") for b in blocks])


class MI355XDenseBackend:
    """Device workspace of the dense Schur route for one problem."""

    name = "mi355x_dense_cholesky"

    def __init__(self, prob: SDPProblem, device: int = 0):
        self.lib = _lib.load()
        self.prob = prob
        self.L, self.m, self.k = prob.L, prob.m, list(prob.k)
        self.k2 = [k * k for k in self.k]
        self.offsets = np.cumsum([0] + self.k2)
        h = C.c_void_p()
        karr = (C.c_int * self.L)(*self.k)
        check(self.lib.sdpx_dense_create(device, self.L, karr, self.m, C.byref(h)), "sdpx_dense_create")
        self.h = h
        for l in range(self.L):
            Cl = np.ascontiguousarray(np.asarray(prob.C[l], dtype=np.float64).ravel(order="F"))
            if l in prob.sparse:
                self._set_block_sparse(l, *prob.sparse[l], Cl)
                continue
            Av = np.ascontiguousarray(prob.Av[l], dtype=np.float64)
            check(self.lib.sdpx_dense_set_block(h, l, dptr(Av), dptr(Cl)), "set_block")
        check(self.lib.sdpx_dense_set_objective(h, dptr(np.ascontiguousarray(prob.c))), "set_objective")
        self._o = np.zeros(4 + 4 * self.L)

    def _set_block_sparse(self, l, ids, coo, Cl):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        arrs = [np.ascontiguousarray(a, dtype=np.int32) for a in (coo.ptr, coo.lin, coo.row, coo.col)]
        val = np.ascontiguousarray(coo.val, dtype=np.float64)
        if arrs[0].size != ids.size + 1 or not (arrs[1].size == arrs[2].size == arrs[3].size == val.size):
            raise ValueError(f"sparse block {l}: ptr needs na+1 entries and lin/row/col/val one per entry")
        ip = _lib._ip
        check(self.lib.sdpx_dense_set_block_sparse(self.h, l, int(ids.size), ids.ctypes.data_as(ip),
                                                   *(a.ctypes.data_as(ip) for a in arrs), dptr(val), dptr(Cl)),
              "set_block_sparse")

    def set_arrow(self, owner=None):
        """Switch the Schur factor and solves to the block-arrow elimination (factor_arrow_kkt!,
        kkt.jl:2132-2349; solve_arrow_kkt!, :2977-3096).  ``owner`` defaults to arrow_owner(prob)."""
        if owner is None:
            owner = arrow_owner(self.prob)
            if owner is None:
                raise ValueError("block_arrow route selected, but the problem is incompatible")
        owner = np.ascontiguousarray(owner, dtype=np.int32)
        if owner.shape != (self.m,):
            raise ValueError(f"set_arrow: owner needs {self.m} entries")
        check(self.lib.sdpx_dense_set_arrow(self.h, owner.ctypes.data_as(_lib._ip)), "set_arrow")

    def arrow_info(self):
        """(global variables or -1, local blocks, largest local block)."""
        v = [C.c_int(0), C.c_int(0), C.c_int(0)]
        check(self.lib.sdpx_dense_get_arrow(self.h, *(C.byref(x) for x in v)), "get_arrow")
        return tuple(x.value for x in v)

    # ---- lifecycle -----------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.sdpx_dense_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- setup on the device ------------------------------------------------
    def equilibrate(self, ruiz_iters: int | str = "auto"):
        """equilibrate (ingest.jl:1432-1521) on the device-resident problem -> Equilibration.

        The handle's panels, C and c are scaled in place; ``self.prob`` is replaced by the scaled
        problem's host copy (C and c downloaded; the panels stay on the device only)."""
        n = -1 if ruiz_iters == "auto" else int(ruiz_iters)
        E = np.empty(int(sum(self.k)))
        s = np.empty(self.m)
        osc = C.c_double(0.0)
        passes = np.zeros(self.L, dtype=np.int32)
        check(self.lib.sdpx_dense_equilibrate(self.h, n, dptr(E), dptr(s), C.byref(osc),
                                              passes.ctypes.data_as(_lib._ip)), "equilibrate")
        Cc = np.empty(int(self.offsets[-1]))
        cc = np.empty(self.m)
        check(self.lib.sdpx_dense_get_data(self.h, dptr(Cc), dptr(cc)), "get_data")
        self.prob = SDPProblem(c=cc, C=self._split(Cc), Av=[None] * self.L, k=list(self.k))
        Es = np.split(E, np.cumsum(self.k)[:-1])
        return Equilibration(E=Es, s=s, objective_scale=float(osc.value), passes=passes.tolist())

    def cold_start(self, beta: float, eps_primal: float, eps_dual: float) -> dict:
        """_kkt_cold_start_initialization (interior_point.jl:699-1287): sets the iterate; returns the record."""
        mu = np.empty(self.L)
        out = np.zeros(16)
        check(self.lib.sdpx_dense_cold_start(self.h, beta, eps_primal, eps_dual, dptr(mu), dptr(out)), "cold_start")
        names = ("ok", "reason", "regularization_attempts", "primal_kkt_residual", "dual_kkt_residual",
                 "refinement_steps", "primal_residual", "dual_residual", "largest_primal_shift",
                 "largest_dual_shift", "primal_margin", "dual_margin", "kappa_after", "kappa_after_mass_floor")
        rec = dict(zip(names, out.tolist()))
        rec["ok"] = bool(rec["ok"])
        rec["reason"] = COLD_START_REASONS[int(rec["reason"])]
        rec["mu"] = mu
        return rec

    # ---- equality-constrained KKT (kkt.jl:1466-1665, 2570-2595) ------------
    def set_equalities(self, B) -> None:
        B = np.asfortranarray(B, dtype=np.float64)
        if B.ndim != 2 or B.shape[0] != self.m:
            raise ValueError(f"B must be m x n_eq with m = {self.m}")
        self.neq = int(B.shape[1])
        check(self.lib.sdpx_dense_set_equalities(self.h, self.neq, dptr(B.ravel(order="F"))), "set_equalities")

    def equality_factor(self, requested: float):
        """-> (status 0 ok / 1 factor failed / 2 rank deficient, min diag, max diag, tau)."""
        out = np.zeros(4)
        rc = self.lib.sdpx_dense_equality_factor(self.h, requested, dptr(out))
        if rc < 0:
            check(rc, "equality_factor")
        return int(out[0]), float(out[1]), float(out[2]), float(out[3])

    def solve_equality(self, p):
        """dx (left on the device; get_direction) and dy for S dx - B dy = r, B' dx = p."""
        dy = np.empty(self.neq)
        check(self.lib.sdpx_dense_solve_equality(self.h, dptr(np.ascontiguousarray(p, dtype=np.float64)), dptr(dy)),
              "solve_equality")
        return dy

    def get_equality(self):
        Bt = np.empty(self.m * self.neq)
        Q = np.empty(self.neq * self.neq)
        sc = np.empty(self.neq)
        check(self.lib.sdpx_dense_get_equality(self.h, dptr(Bt), dptr(Q), dptr(sc)), "get_equality")
        return Bt.reshape(self.m, self.neq, order="F"), Q.reshape(self.neq, self.neq, order="F"), sc

    def memory_bytes(self) -> int:
        b = C.c_longlong()
        check(self.lib.sdpx_dense_memory(self.h, C.byref(b)), "memory")
        return int(b.value)

    # ---- iterate I/O -----------------------------------------------------
    def set_iterate(self, x, X, Y):
        xa = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        Xa = None if X is None else _fortran_concat(X)
        Ya = None if Y is None else _fortran_concat(Y)
        check(self.lib.sdpx_dense_set_iterate(self.h, dptr(xa), dptr(Xa), dptr(Ya)), "set_iterate")

    def _split(self, flat):
        return [flat[self.offsets[l]:self.offsets[l + 1]].reshape(self.k[l], self.k[l], order="F").copy()
                for l in range(self.L)]

    def get_iterate(self):
        x = np.empty(self.m)
        X = np.empty(int(self.offsets[-1]))
        Y = np.empty(int(self.offsets[-1]))
        check(self.lib.sdpx_dense_get_iterate(self.h, dptr(x), dptr(X), dptr(Y)), "get_iterate")
        return x, self._split(X), self._split(Y)

    def get_direction(self):
        dx = np.empty(self.m)
        dX = np.empty(int(self.offsets[-1]))
        dY = np.empty(int(self.offsets[-1]))
        check(self.lib.sdpx_dense_get_direction(self.h, dptr(dx), dptr(dX), dptr(dY)), "get_direction")
        return dx, self._split(dX), self._split(dY)

    def get_schur(self) -> np.ndarray:
        """m x m Schur matrix; only the lower triangle is authoritative."""
        S = np.empty((self.m, self.m), order="F")
        check(self.lib.sdpx_dense_get_schur(self.h, dptr(S.ravel(order="K"))), "get_schur")
        return S

    def set_schur(self, S):
        Sf = np.asfortranarray(S, dtype=np.float64)
        check(self.lib.sdpx_dense_set_schur(self.h, dptr(Sf.ravel(order="K"))), "set_schur")

    def set_rhs(self, r):
        check(self.lib.sdpx_dense_set_rhs(self.h, dptr(np.ascontiguousarray(r, dtype=np.float64))), "set_rhs")

    def get_rhs(self):
        r = np.empty(self.m)
        check(self.lib.sdpx_dense_get_rhs(self.h, dptr(r)), "get_rhs")
        return r

    def set_dx(self, dx):
        check(self.lib.sdpx_dense_set_dx(self.h, dptr(np.ascontiguousarray(dx, dtype=np.float64))), "set_dx")

    # ---- Newton-step phases -------------------------------------------------
    def compute_residuals(self, mu, affine: bool, factor: bool):
        """(p_res, d_res, blocks_ok) -- threaded_compute_residuals! (threaded.jl:423-505)."""
        mu_a = np.ascontiguousarray(mu, dtype=np.float64)
        out = np.zeros(3)
        s = check(self.lib.sdpx_dense_residuals(self.h, dptr(mu_a), int(affine), int(factor), dptr(out)), "residuals")
        return float(out[0]), float(out[1]), s == 0

    def schur_build(self):
        """threaded_schur_build! -> schur_build! (schur.jl:511-547)."""
        check(self.lib.sdpx_dense_schur_build(self.h), "schur_build")

    def factorize(self):
        """factorize!(::DenseCholeskyBackend) (kkt_backend.jl:211-221) -> (ok, reg_attempts)."""
        a = C.c_int(0)
        s = check(self.lib.sdpx_dense_schur_factor(self.h, C.byref(a)), "schur_factor")
        return s == 0, int(a.value)

    def predictor_rhs(self):
        check(self.lib.sdpx_dense_predictor_rhs(self.h), "predictor_rhs")

    def corrector_rhs(self, mehrotra: bool, sigma: float, mu_global: float, block_local: bool, mu, beta: float):
        mu_a = np.ascontiguousarray(mu, dtype=np.float64)
        check(self.lib.sdpx_dense_corrector_rhs(self.h, int(mehrotra), sigma, mu_global, int(block_local),
                                                dptr(mu_a), beta), "corrector_rhs")

    def solve_direction(self) -> bool:
        """solve_direction! (kkt_backend.jl:392-410)."""
        check(self.lib.sdpx_dense_solve(self.h), "solve")
        return True

    def direction_residual(self):
        out = np.zeros(2)
        check(self.lib.sdpx_dense_direction_residual(self.h, dptr(out)), "direction_residual")
        return float(out[0]), float(out[1])

    def refine(self, abstol: float, max_steps: int, fixed_policy: bool):
        """refine! (kkt_backend.jl:476-485) -> (steps, residual)."""
        steps = C.c_int(0)
        res = C.c_double(0.0)
        check(self.lib.sdpx_dense_refine(self.h, abstol, int(max_steps), int(fixed_policy), C.byref(steps),
                                         C.byref(res)), "refine")
        return int(steps.value), float(res.value)

    def direction_blocks(self):
        check(self.lib.sdpx_dense_direction_blocks(self.h), "direction_blocks")

    def complementarity(self, tX: float, tY: float):
        out = np.zeros(2 * self.L)
        check(self.lib.sdpx_dense_complementarity(self.h, tX, tY, dptr(out)), "complementarity")
        return out[: self.L].copy(), out[self.L:].copy()

    def fraction_to_boundary(self, pf: float, df: float):
        tx, ty = C.c_double(), C.c_double()
        check(self.lib.sdpx_dense_fraction_to_boundary(self.h, pf, df, C.byref(tx), C.byref(ty)), "ftb")
        return float(tx.value), float(ty.value)

    def backtrack(self, gamma: float, min_step: float, ratio: float):
        tx, ty = C.c_double(), C.c_double()
        check(self.lib.sdpx_dense_backtrack(self.h, gamma, min_step, ratio, C.byref(tx), C.byref(ty)), "backtrack")
        return float(tx.value), float(ty.value)

    def update(self, tX: float, tY: float):
        out = np.zeros(2 + self.L)
        check(self.lib.sdpx_dense_update(self.h, tX, tY, dptr(out)), "update")
        return float(out[0]), bool(out[1] != 0.0), out[2:].copy()

    def accepted_residuals(self, tX: float, tY: float, block_primal_residual: float):
        out = np.zeros(2)
        check(self.lib.sdpx_dense_accepted_residuals(self.h, tX, tY, block_primal_residual, dptr(out)),
              "accepted_residuals")
        return float(out[0]), float(out[1])

    def objectives(self):
        p, d = C.c_double(), C.c_double()
        check(self.lib.sdpx_dense_objectives(self.h, C.byref(p), C.byref(d)), "objectives")
        return float(p.value), float(d.value)

    def solution_residuals(self):
        p, d = C.c_double(), C.c_double()
        check(self.lib.sdpx_dense_solution_residuals(self.h, C.byref(p), C.byref(d)), "solution_residuals")
        return float(p.value), float(d.value)

    def scale_blocks(self, which: int, factor: float):
        check(self.lib.sdpx_dense_scale_blocks(self.h, which, factor), "scale_blocks")

    def factor_condition(self, which: int = 3):
        """(cond_1(S) estimate as xPOCON, the xTRCON product of _triangular_condition_estimate
        (mixed_precision_kkt.jl:371-392), rcond_1(L), rcond_inf(L)) from the last Schur factor;
        which = 1 computes the first only, 2 the others only."""
        out = np.zeros(4)
        check(self.lib.sdpx_dense_factor_condition(self.h, int(which), dptr(out)), "factor_condition")
        return tuple(float(v) for v in out)

    def factor_quality(self):
        out = np.zeros(3)
        check(self.lib.sdpx_dense_factor_quality(self.h, dptr(out)), "factor_quality")
        return float(out[0]), float(out[1]), float(out[2])

    def set_profiling(self, enabled: bool):
        check(self.lib.sdpx_dense_set_profiling(self.h, int(enabled)), "set_profiling")

    def kernel_times(self, reset: bool = False) -> dict:
        """Device-event milliseconds of the Schur Gram kernel, panels, reduction and factor."""
        out = np.zeros(6)
        check(self.lib.sdpx_dense_kernel_times(self.h, int(reset), dptr(out)), "kernel_times")
        return dict(gram_ms=out[0], gram_launches=int(out[1]), panels_ms=out[2], reduce_ms=out[3],
                    factor_ms=out[4], factor_calls=int(out[5]))

    def autotune(self) -> dict:
        """Time the Schur build's tiling candidates on the device and keep the fastest (plan time)."""
        out = np.zeros(20)
        check(self.lib.sdpx_dense_autotune(self.h, dptr(out)), "autotune")
        shapes = ("128x128", "64x64", "128x64", "64x128")
        return dict(panel_tiles=[shapes[int(out[0])], shapes[int(out[1])]], split=int(out[2]),
                    tile_group=int(out[3]), panel_ms=round(float(out[4]), 4), gram_ms=round(float(out[5]), 4),
                    default_split=int(out[6]),
                    panel_candidates_ms=[{sh: round(float(out[7 + 4 * p + c]), 4) for c, sh in enumerate(shapes)}
                                         for p in range(2)],
                    ftb_levels_per_round=int(out[15]),
                    ftb_model_ms={d: round(float(out[16 + d - 4]), 4) for d in range(4, 8)})

    def set_tiling(self, panel_tiles=("64x64", "64x64"), split: int = 1, tile_group: int = 8,
                   ftb_levels_per_round: int = 7) -> None:
        """Apply tiling choices (e.g. a saved autotune result) without timing them."""
        shapes = ("128x128", "64x64", "128x64", "64x128")
        check(self.lib.sdpx_dense_set_tiling(self.h, shapes.index(panel_tiles[0]), shapes.index(panel_tiles[1]),
                                             int(split), int(tile_group), int(ftb_levels_per_round)), "set_tiling")

    def schur_split(self) -> int:
        s = C.c_int(0)
        check(self.lib.sdpx_dense_schur_split(self.h, C.byref(s)), "schur_split")
        return int(s.value)

    # ---- mixed precision -------------------------------------------------------------
    def set_factor_precision(self, precision: int):
        """Schur factor in FP64 (0) or FP32 (1: FP32 MFMA panels and trailing updates)."""
        check(self.lib.sdpx_dense_set_factor_precision(self.h, int(precision)), "set_factor_precision")

    def set_panel_precision(self, precision: int):
        """P-hat products of the BF16x3 assembly in FP64 (0) or on FP32 MFMA (1)."""
        check(self.lib.sdpx_dense_set_panel_precision(self.h, int(precision)), "set_panel_precision")

    def set_precision(self, precision: int):
        """0: FP64 Schur assembly; 1: FP32 Schur Gram product; 2: BF16x3 Schur Gram product (S kept in FP64)."""
        check(self.lib.sdpx_dense_set_precision(self.h, int(precision)), "set_precision")

    def exact_residual(self):
        """(||r - S dx||_inf with S applied matrix-free in FP64, ||r||_inf)."""
        out = np.zeros(2)
        check(self.lib.sdpx_dense_exact_residual(self.h, dptr(out)), "exact_residual")
        return float(out[0]), float(out[1])

    def refine_exact(self, abstol: float, max_steps: int):
        """Refinement against the matrix-free FP64 operator -> (steps, residual)."""
        steps = C.c_int(0)
        res = C.c_double(0.0)
        check(self.lib.sdpx_dense_refine_exact(self.h, abstol, int(max_steps), C.byref(steps), C.byref(res)),
              "refine_exact")
        return int(steps.value), float(res.value)

    def synchronize(self):
        check(self.lib.sdpx_dense_synchronize(self.h), "synchronize")
