"""CPU Float64 oracle for the SDPX dense Newton step — TEST INFRASTRUCTURE ONLY.

This module restates, in numpy/scipy (OpenBLAS/LAPACK), the reference
algorithm of SDPX.jl's dense primal-dual interior-point path: residuals,
Schur assembly, Schur Cholesky with the regularisation ladder, triangular
solves, iterative refinement, predictor/corrector right-hand sides, direction
recovery, step length / PSD checks, the accepted-step update, the fixed and
adaptive parameter controllers, stagnation detection and the IPM loop.

It is the checker for the HIP path. Only ``tests/``, ``__graft_entry__.smoke``
and ``bench.py``'s ``cpu_baseline`` leg may import it; no product code path may
call it.

Every function cites the reference file:line it follows (paths relative to the
SDPX.jl repository root).  Parity pins (see ``tests/test_oracle.py``):

* ``kchol``/``factor_dense_kkt`` -- LAPACK dpotrf, closed forms for k<=2, and
  the reference's regularisation contract (kkt_regressions.jl:927-1000).
* ``schur_build`` -- brute-force tr(A_i X^-1 A_j Y) in 50-digit mpmath and the
  X=Y=I Gram closed form.
* ``solve_dense_kkt``/``refine_direction`` -- S dx = r identities.
* ``factor_arrow_kkt``/``solve_arrow_kkt`` -- the dense solve of the same S (the elimination
  is exact), the reference's three-block case (test/sparse.jl:8-26, 141-188) and its
  at-scale arrow-vs-dense check (test/sparse.jl:437-495) on the same structure, the
  per-block ladder on a planted indefinite local block.
* ``sparse_schur_block``/``scatter_sparse_schur`` -- the dense Gram route on the same
  blocks (the identity tr(A_p X^-1 A_r Y)), the reference's own "general sparse
  outer-product Schur equals dense path" case (test/sparse.jl:191-258) and its analytic
  sparse solve (test/sparse.jl:330-352, p* = 2 sqrt 6).
* ``fraction_to_boundary_bound`` -- eigenvalue bound (scipy eigh) and the 2x2
  determinant closed form.
* ``solve`` -- reference known answers T1 (p* = 2 sqrt 6, 19 +- 2 iterations,
  test/correctness.jl:71-94, 230-243) and strong duality of the seeded
  feasible generator.
* ``AdaptivePolicy``/``Controller`` -- the reference's own known values of
  test/adaptive_parameter_policy.jl:20-285 (sigma = 0.15 under the cap, 3/25,
  fallback reasons and the restored fixed policy, achieved reduction 0.75),
  checked on this oracle and on both product controllers
  (tests/test_adaptive_policy.py).  The whole-solve trajectory stays unpinned
  against the reference (Julia is unavailable); near the gap target some
  statuses depend on the summation order (``summation_order``,
  tests/test_status_rounding.py).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np
import scipy.linalg as sla
from scipy.linalg import lapack

EPS = float(np.finfo(np.float64).eps)
SQRT_EPS = math.sqrt(EPS)

# kkt.jl:3634-3637
REFINE_MIN_DECREASE = 0.5
REFINE_DEFAULT_TOL_ULPS = 64
# interior_point.jl:2703-2709
NEAR_SOLUTION_MERIT = 100.0
CENTERING_BETA_STEP = 4.0
CENTERING_BETA_MAX = 0.5
# stagnation.jl:13, 26
PRECISION_FLOOR_ULPS = 16
PROJECTION_SLACK = 10.0


# --------------------------------------------------------------------------
# Problem data (types/problems.jl:1-12, types/constraints.jl:26-28)
# --------------------------------------------------------------------------
@dataclass
class Problem:
    """Core geometric SDP: min c'x s.t. X_l = sum_i x_i A_i^(l) - C_l >= 0.

    ``Av[l]`` is the k_l^2 x m column-major constraint panel of the reference's
    ``DenseCons`` (ingest.jl:680-694): column i is vec(A_i^(l)).
    """

    c: np.ndarray
    C: list
    Av: list
    k: list
    # blocks assembled by the sparse Schur route: l -> (ids, SparseBlockCOO); Av[l] stays the
    # dense panel the residual and direction passes read
    sparse: dict = field(default_factory=dict)

    @property
    def m(self) -> int:
        return int(self.c.shape[0])

    @property
    def L(self) -> int:
        return len(self.k)


def ingest(c, A, C) -> Problem:
    """Dense ingest of ``A[l]`` given as (m, k, k) arrays (ingest.jl:680-694).

    The reference's T1 data (test/correctness.jl:71-85) uses exactly this
    (m, k, k) layout, A[l][i, :, :] = A_i^(l).
    """
    c = np.asarray(c, dtype=np.float64).copy()
    Av, Cs, ks = [], [], []
    for Al, Cl in zip(A, C):
        Al = np.asarray(Al, dtype=np.float64)
        m, k, _ = Al.shape
        # vec() is column-major; A_i is symmetric so A_i.T.ravel() == A_i.ravel()
        Av.append(np.ascontiguousarray(Al.transpose(0, 2, 1).reshape(m, k * k).T))
        Cs.append(np.asarray(Cl, dtype=np.float64).copy())
        ks.append(k)
    return Problem(c=c, C=Cs, Av=Av, k=ks)


# --------------------------------------------------------------------------
# Options (types/workspaces.jl:101-228) and controller policies
# (adaptive_parameters.jl:21-186)
# --------------------------------------------------------------------------
@dataclass
class Options:
    beta: float = 0.1
    gamma: float = 0.9
    omega_p: float = 1.0
    omega_d: float = 1.0
    eps_gap: float = 1e-10
    eps_primal: float = 1e-10
    eps_dual: float = 1e-10
    iter_max: int = 200
    restart: bool = True
    min_step: float = 1e-10
    omega_step: float = 1e5
    max_restarts: int = 5
    max_centering: int = 4
    stall_iterations: int = 15
    stall_tolerance: float = 1e-3
    step_rule: str = "auto"
    predictor: str = "classic"
    refine_steps: int = 1
    refine_policy: str = "auto"
    refine_max_steps: int = 8
    refine_tol: float = 0.0
    parameter_strategy: str = "adaptive"
    adaptive_sigma_max: float = 0.0
    # KKT route of the Schur system: "dense_cholesky" or "block_arrow" (the exact
    # block-arrow elimination, kkt.jl:2132-2349, 2977-3096)
    kkt_route: str = "dense_cholesky"


@dataclass
class IterationParameters:
    """adaptive_parameters.jl:95-110"""

    sigma: float
    primal_fraction_to_boundary: float
    dual_fraction_to_boundary: float
    backtracking_factor: float
    refinement_tolerance: float
    refinement_max_count: int
    minimum_step: float
    restart_scale: float
    primal_regularization: float = 0.0
    dual_regularization: float = 0.0
    fallback: bool = False
    fallback_reason: str = "none"


@dataclass
class FixedPolicy:
    """adaptive_parameters.jl:21-32, 112-131"""

    sigma: float
    primal_fraction: float
    dual_fraction: float
    backtracking: float
    refinement_tolerance: float
    refinement_max_count: int
    minimum_step: float
    restart_scale: float

    @staticmethod
    def from_options(o: Options) -> "FixedPolicy":
        tol = o.refine_tol if o.refine_tol > 0 else REFINE_DEFAULT_TOL_ULPS * EPS
        count = o.refine_steps if o.refine_policy == "fixed" else o.refine_max_steps
        return FixedPolicy(o.beta, o.gamma, o.gamma, o.gamma, tol, count, o.min_step, o.omega_step)

    def params(self, fallback=False, reason="none") -> IterationParameters:
        """adaptive_parameters.jl:160-186"""
        return IterationParameters(
            sigma=self.sigma,
            primal_fraction_to_boundary=self.primal_fraction,
            dual_fraction_to_boundary=self.dual_fraction,
            backtracking_factor=self.backtracking,
            refinement_tolerance=self.refinement_tolerance,
            refinement_max_count=self.refinement_max_count,
            minimum_step=self.minimum_step,
            restart_scale=self.restart_scale,
            fallback=fallback,
            fallback_reason=reason,
        )


@dataclass
class AdaptivePolicy:
    """adaptive_parameters.jl:40-53, 133-158"""

    fallback: FixedPolicy
    sigma_min: float
    sigma_max: float
    fraction_min: float
    fraction_max: float
    backtracking_min: float
    backtracking_max: float
    centrality_floor: float
    severe_centrality_floor: float
    refinement_floor: float
    requested_tolerance: float
    instability_limit: int

    @staticmethod
    def from_options(o: Options) -> "AdaptivePolicy":
        fixed = FixedPolicy.from_options(o)
        requested = min(o.eps_gap, o.eps_primal, o.eps_dual)
        refinement_floor = max(REFINE_DEFAULT_TOL_ULPS * EPS, requested * requested if requested > 0 else 0.0)
        general = max(0.5, fixed.sigma)
        sigma_max = max(o.adaptive_sigma_max, fixed.sigma) if o.adaptive_sigma_max > 0 else general
        return AdaptivePolicy(
            fixed,
            min(1 / 50, fixed.sigma),
            sigma_max,
            min(4 / 5, fixed.primal_fraction),
            max(99 / 100, fixed.primal_fraction),
            min(11 / 20, fixed.backtracking),
            max(9 / 10, fixed.backtracking),
            2 / 25,
            1 / 5,
            refinement_floor,
            requested,
            2,
        )


@dataclass
class Diagnostics:
    """adaptive_parameters.jl:63-83 (IterationDiagnostics)."""

    iteration: int = 0
    primal_residual: float = math.inf
    dual_residual: float = math.inf
    relative_gap: float = math.inf
    mu: float = math.inf
    mu_aff: float = math.inf
    affine_primal_step: float = 0.0
    affine_dual_step: float = 0.0
    previous_primal_step: float = 1.0
    previous_dual_step: float = 1.0
    backtracking_count: int = 0
    regularization: float = 0.0
    refinement_count: int = 0
    factorization_quality: float = 1.0
    achieved_residual_reduction: float = 1.0
    primal_psd_margin: float = 1.0
    dual_psd_margin: float = 1.0


def _finite_diagnostics(d: Diagnostics) -> bool:
    """adaptive_parameters.jl:211-226"""
    vals = (d.primal_residual, d.dual_residual, d.mu, d.mu_aff, d.affine_primal_step,
            d.affine_dual_step, d.factorization_quality, d.primal_psd_margin, d.dual_psd_margin)
    return all(math.isfinite(v) for v in vals) and d.mu > 0 and d.mu_aff >= 0


def _adaptive_fraction(p: AdaptivePolicy, affine_step, previous_step, backtracking_count, ratio):
    """adaptive_parameters.jl:228-251"""
    if affine_step < 1 / 20:
        cand = 21 / 25
    elif affine_step < 1 / 5:
        cand = 9 / 10
    elif affine_step < 7 / 10:
        cand = 19 / 20
    else:
        cand = 99 / 100
    if previous_step < 1 / 10:
        cand = min(cand, 9 / 10)
    if backtracking_count >= 4:
        cand = min(cand, 22 / 25)
    if ratio < 1 / 1_000_000:
        cand = min(cand, 9 / 10)
    return min(max(cand, p.fraction_min), p.fraction_max)


def _recent_instability_count(history) -> int:
    """adaptive_parameters.jl:202-209"""
    count = 0
    for row in reversed(history):
        if not row.get("unstable", False):
            break
        count += 1
    return count


def _history_value(history, name, default):
    """adaptive_parameters.jl:196-200"""
    if not history:
        return default
    return history[-1].get(name, default)


def select_adaptive_parameters(p: AdaptivePolicy, d: Diagnostics, history) -> IterationParameters:
    """adaptive_parameters.jl:253-453"""
    if not _finite_diagnostics(d):
        return p.fallback.params(True, "nonfinite_diagnostics")
    if d.factorization_quality <= 0:
        return p.fallback.params(True, "degraded_factorization")
    if _recent_instability_count(history) >= p.instability_limit:
        return p.fallback.params(True, "repeated_instability")
    cold_merit = max(d.primal_residual, d.dual_residual, d.relative_gap)
    if cold_merit >= 1:
        return p.fallback.params()
    ratio = min(max(d.mu_aff / d.mu, 0.0), 1.0)
    affine_min = min(d.affine_primal_step, d.affine_dual_step)
    exponent = 3 if affine_min >= 3 / 4 else 2
    sigma_candidate = ratio ** exponent
    if affine_min < 1 / 20 or d.achieved_residual_reduction > 5 / 4 or d.factorization_quality < SQRT_EPS:
        floor = p.severe_centrality_floor
    elif (affine_min < 1 / 4 or d.backtracking_count >= 4
          or d.refinement_count >= max(p.fallback.refinement_max_count - 1, 1)):
        floor = p.centrality_floor
    else:
        floor = 0.0
    sigma_candidate = max(sigma_candidate, floor)
    if d.iteration <= 1:
        sigma_candidate = max(sigma_candidate, min(p.fallback.sigma, p.sigma_max))
    previous_sigma = _history_value(history, "sigma", p.fallback.sigma)
    sigma = min(max(4 / 5 * sigma_candidate + 1 / 5 * previous_sigma, p.sigma_min), p.sigma_max)
    pf = _adaptive_fraction(p, d.affine_primal_step, d.previous_primal_step, d.backtracking_count, ratio)
    df = _adaptive_fraction(p, d.affine_dual_step, d.previous_dual_step, d.backtracking_count, ratio)
    if d.backtracking_count >= 6 or d.achieved_residual_reduction > 5 / 4:
        bt = 13 / 20
    elif d.backtracking_count >= 3:
        bt = 18 / 25
    elif affine_min > 4 / 5:
        bt = 17 / 20
    else:
        bt = 4 / 5
    bt = min(max(bt, p.backtracking_min), p.backtracking_max)
    reg = max(d.regularization / 10, SQRT_EPS) if d.regularization > 0 else 0.0
    refinement_tolerance = max(p.refinement_floor, p.fallback.refinement_tolerance)
    if d.factorization_quality < SQRT_EPS or d.refinement_count > 0:
        refinement_max_count = p.fallback.refinement_max_count
    else:
        refinement_max_count = min(p.fallback.refinement_max_count, 2)
    requested = p.requested_tolerance if p.requested_tolerance > 0 else SQRT_EPS
    precision_step_floor = max(100 * EPS, min(p.fallback.minimum_step, requested / 100))
    if d.factorization_quality < SQRT_EPS or d.backtracking_count >= 8:
        minimum_step = p.fallback.minimum_step
    else:
        minimum_step = precision_step_floor
    fmin = max(min(d.primal_residual, d.dual_residual), requested)
    fratio = max(d.primal_residual, d.dual_residual) / fmin
    cand = 1000.0 if fratio > 1e6 else (100.0 if fratio > 1000 else 10.0)
    restart_scale = min(max(cand, 10.0), p.fallback.restart_scale)
    return IterationParameters(
        sigma=sigma,
        primal_fraction_to_boundary=pf,
        dual_fraction_to_boundary=df,
        backtracking_factor=bt,
        refinement_tolerance=refinement_tolerance,
        refinement_max_count=refinement_max_count,
        minimum_step=minimum_step,
        restart_scale=restart_scale,
        primal_regularization=reg,
        dual_regularization=reg,
    )


class Controller:
    """AdaptiveIPMController (adaptive_parameters.jl:440-476, 510-548, 597-769)."""

    def __init__(self, o: Options):
        self.strategy = o.parameter_strategy
        self.default_beta = o.beta
        self.default_gamma = o.gamma
        self.beta = o.beta
        self.gamma = o.gamma
        self.previous_feasibility = math.nan
        self.instability_count = 0
        self.fallback = False
        self.history: list[dict] = []
        self.policy = AdaptivePolicy.from_options(o) if o.parameter_strategy == "adaptive" else FixedPolicy.from_options(o)
        self.parameters = FixedPolicy.from_options(o).params()

    def select(self, d: Diagnostics) -> IterationParameters:
        """adaptive_parameters.jl:510-548"""
        if self.fallback:
            fp = FixedPolicy(self.default_beta, self.default_gamma, self.default_gamma, self.default_gamma,
                             self.parameters.refinement_tolerance, self.parameters.refinement_max_count,
                             self.parameters.minimum_step, self.parameters.restart_scale)
            params = fp.params(True, "previous_instability")
        elif isinstance(self.policy, FixedPolicy):
            params = self.policy.params()
        else:
            params = select_adaptive_parameters(self.policy, d, self.history)
        self.parameters = params
        self.beta = params.sigma
        self.gamma = min(params.primal_fraction_to_boundary, params.dual_fraction_to_boundary)
        if params.fallback:
            self.fallback = True
            self.beta = self.default_beta
            self.gamma = self.default_gamma
        return params

    def record(self, *, iteration, complementarity_before, complementarity_after, primal_residual,
               dual_residual, primal_step, dual_step, primal_residual_after, dual_residual_after,
               backtracking_count, refinement_count, selected: IterationParameters):
        """adaptive_parameters.jl:597-769 (fields consumed by the policies)."""
        reduction = complementarity_after / complementarity_before if complementarity_before > 0 else 1.0
        feasibility = max(primal_residual, dual_residual)
        prev = self.previous_feasibility
        fratio = feasibility / prev if (math.isfinite(prev) and prev > 0) else 1.0
        unstable, reason = _outcome_is_unstable(iteration, reduction, fratio, primal_step, dual_step)
        if self.strategy == "adaptive" and not self.fallback:
            severe = reason in ("nonfinite_outcome", "complementarity_growth", "feasibility_growth")
            if severe:
                self.instability_count = 2
            elif unstable:
                self.instability_count += 1
            else:
                self.instability_count = max(self.instability_count - 1, 0)
            if self.instability_count >= 2:
                self.beta = self.default_beta
                self.gamma = self.default_gamma
                self.fallback = True
        residual_before = max(primal_residual, dual_residual, EPS)
        achieved = max(primal_residual_after, dual_residual_after) / residual_before
        self.history.append(dict(
            iteration=iteration, sigma=selected.sigma, primal_step=primal_step, dual_step=dual_step,
            backtracking_count=backtracking_count, refinement_count=refinement_count,
            primal_residual=primal_residual, dual_residual=dual_residual,
            achieved_residual_reduction=achieved, unstable=unstable,
        ))
        self.previous_feasibility = feasibility

    def iteration_options(self, o: Options) -> Options:
        """controller_options (adaptive_parameters.jl:771-782)."""
        return replace(o, beta=self.beta, gamma=self.gamma,
                       parameter_strategy="fixed" if self.fallback else self.strategy)


def _outcome_is_unstable(iteration, reduction, fratio, primal_step, dual_step):
    """adaptive_parameters.jl:550-586"""
    nonfinite = not all(math.isfinite(v) for v in (reduction, fratio, primal_step, dual_step))
    comp_growth = iteration > 3 and reduction > 2
    feas_growth = iteration > 3 and fratio > 3
    min_step = min(primal_step, dual_step)
    tiny = iteration > 3 and primal_step < 1e-6 and dual_step < 1e-6 and fratio >= 1
    stalled = iteration > 5 and min_step < 1 / 50 and fratio > 49 / 50 and reduction > 49 / 50
    reason = ("nonfinite_outcome" if nonfinite else "complementarity_growth" if comp_growth else
              "feasibility_growth" if feas_growth else "repeated_tiny_step" if tiny else
              "stalled_progress" if stalled else "none")
    return reason != "none", reason


def estimate_backtracking_count(step, gamma, rule):
    """adaptive_parameters.jl:500-508"""
    if rule != "backtrack":
        return 0
    if not math.isfinite(step) or step <= 0 or step >= 1 or gamma <= 0 or gamma >= 1:
        return 0
    ratio = math.log(step) / math.log(gamma)
    return max(0, int(round(ratio)))


class Stagnation:
    """StagnationDetector (stagnation.jl:56-177)."""

    def __init__(self, window, tolerance):
        self.window = max(window, 0)
        self.tolerance = max(tolerance, 0.0)
        self.merits: list[float] = []
        self.at_floor: list[bool] = []
        self.reason = "none"

    @staticmethod
    def merit(o: Options, p_res, d_res, gap_rel, comp, scale_p, scale_d, objective_scale):
        """stagnation.jl:89-100"""
        safe = lambda e: e if e > 0 else EPS  # noqa: E731
        merit = (p_res / scale_p) / safe(o.eps_primal)
        merit = max(merit, (d_res / scale_d) / safe(o.eps_dual))
        merit = max(merit, gap_rel / safe(o.eps_gap))
        if math.isfinite(comp) and comp > 0:
            merit = max(merit, comp / (safe(o.eps_gap) * max(1.0, objective_scale)))
        return merit

    @staticmethod
    def at_precision_floor(p_res, d_res, gap_rel, scale_p, scale_d):
        """stagnation.jl:108-113"""
        f = PRECISION_FLOOR_ULPS * EPS
        return p_res / scale_p <= f and d_res / scale_d <= f and gap_rel <= f

    def observe(self, merit, floor_reached, remaining) -> bool:
        """stagnation.jl:131-167"""
        if self.window <= 0:
            return False
        self.merits.append(merit)
        self.at_floor.append(floor_reached)
        if len(self.merits) > self.window + 1:
            self.merits.pop(0)
            self.at_floor.pop(0)
        if len(self.merits) <= self.window:
            return False
        ref = self.merits[0]
        if not (math.isfinite(ref) and math.isfinite(merit) and ref > 0 and merit > 0):
            return False
        progress = math.log(ref) - math.log(merit)
        rate = progress / self.window
        if not progress > self.tolerance:
            self.reason = "precision_floor" if all(self.at_floor) else "no_progress"
            return True
        if merit <= 1:
            return False
        projected = math.log(merit) / rate
        if projected > PROJECTION_SLACK * max(remaining, 1):
            self.reason = "precision_floor" if all(self.at_floor) else "too_slow"
            return True
        return False


# --------------------------------------------------------------------------
# Kernel primitives (kernels/generic.jl)
# --------------------------------------------------------------------------
def knrm_inf(a) -> float:
    """generic.jl:176"""
    a = np.asarray(a)
    return 0.0 if a.size == 0 else float(np.max(np.abs(a)))


def kchol(A: np.ndarray):
    """In-place-free lower Cholesky without throwing (generic.jl:133-154).

    Returns (L, ok); L is lower triangular with zeros above the diagonal.
    """
    k = A.shape[0]
    if k == 1:
        d11 = A[0, 0]
        if not d11 > 0:
            return None, False
        return np.array([[math.sqrt(d11)]]), True
    if k == 2:
        d11 = A[0, 0]
        if not d11 > 0:
            return None, False
        l11 = math.sqrt(d11)
        l21 = A[1, 0] / l11
        d22 = A[1, 1] - l21 * l21
        if not d22 > 0:
            return None, False
        return np.array([[l11, 0.0], [l21, math.sqrt(d22)]]), True
    c, info = lapack.dpotrf(np.asfortranarray(A), lower=1, clean=1)
    if info != 0:
        return None, False
    return np.tril(c), True


def kcholsolve(LX: np.ndarray, M: np.ndarray) -> np.ndarray:
    """X^-1 M from the lower factor of X (generic.jl:196-242)."""
    k = LX.shape[0]
    if k == 2:
        l11, l21, l22 = LX[0, 0], LX[1, 0], LX[1, 1]
        out = np.empty_like(M)
        y1 = M[0] / l11
        y2 = (M[1] - l21 * y1) / l22
        x2 = y2 / l22
        out[0] = (y1 - l21 * x2) / l11
        out[1] = x2
        return out
    if k == 1:
        return M / LX[0, 0] / LX[0, 0]
    t = sla.solve_triangular(LX, M, lower=True, check_finite=False)
    return sla.solve_triangular(LX, t, lower=True, trans="T", check_finite=False)


def trial_isposdef(X, t, dX) -> bool:
    """generic.jl:292-305"""
    k = X.shape[0]
    if k == 1:
        return X[0, 0] + t * dX[0, 0] > 0
    if k == 2:
        a = X[0, 0] + t * dX[0, 0]
        if not a > 0:
            return False
        b = X[1, 0] + t * dX[1, 0]
        d = X[1, 1] + t * dX[1, 1]
        return a * d > b * b
    return kchol(X + t * dX)[1]


def trial_has_cholesky_margin(X, t, dX, minimum_ratio) -> bool:
    """generic.jl:318-340"""
    if not minimum_ratio > 0:
        return trial_isposdef(X, t, dX)
    Lt, ok = kchol(X + t * dX)
    if not ok:
        return False
    diag = np.abs(np.diag(Lt))
    mn, mx = float(diag.min()), float(diag.max())
    return math.isfinite(mn) and math.isfinite(mx) and mx > 0 and mn > minimum_ratio * mx


def fraction_to_boundary_bound(X, dX) -> float:
    """Largest t in [0,1] keeping X + t dX PD (generic.jl:350-420)."""
    k = X.shape[0]
    if k == 1:
        d = dX[0, 0]
        return min(1.0, -X[0, 0] / d) if d < 0 else 1.0
    if k == 2:
        a, b, d = X[0, 0], X[0, 1], X[1, 1]
        da, db, dd = dX[0, 0], (dX[0, 1] + dX[1, 0]) / 2.0, dX[1, 1]
        a1, b1, d1 = a + da, b + db, d + dd
        if a1 > 0 and a1 * d1 > b1 * b1:
            return 1.0
        c0 = a * d - b * b
        c1 = da * d + a * dd - 2.0 * b * db
        c2 = da * dd - db * db
        root, found = 1.0, False
        if c2 == 0:
            if c1 < 0:
                cand = -c0 / c1
                if 0 < cand <= 1:
                    root, found = cand, True
        else:
            disc = c1 * c1 - 4.0 * c2 * c0
            if disc >= 0:
                sd = math.sqrt(disc)
                den = 2.0 * c2
                c1a = (-c1 - sd) / den
                c2a = (-c1 + sd) / den
                if 0 < c1a <= root:
                    root, found = c1a, True
                if 0 < c2a <= root:
                    root, found = c2a, True
        if found:
            return root
    elif trial_isposdef(X, 1.0, dX):
        return 1.0
    lower, upper = 0.0, 1.0
    for _ in range(64):
        mid = (lower + upper) / 2.0
        if trial_isposdef(X, mid, dX):
            lower = mid
        else:
            upper = mid
    return lower


def symmetrize(M):
    """generic.jl:425-433"""
    return (M + M.T) / 2.0


# --------------------------------------------------------------------------
# Constraint contractions and Schur assembly (schur.jl)
# --------------------------------------------------------------------------
# Summation order of the two A-contractions below.  "blas" is OpenBLAS dgemv,
# the reference's own call (schur.jl:36-44 through LinearAlgebra).  The others
# are equally valid Float64 evaluations of the same sums, used by the tests to
# show which termination statuses are decided by rounding: "pairwise" is
# numpy's pairwise reduction over the summed index, "strided:N" adds N
# interleaved partial GEMVs (terms j = s mod N) in order s = 0..N-1.
_ORDER = {"contract": "blas", "build": "blas"}


class summation_order:
    """Context manager selecting the order of ``contract`` / ``build_P`` sums."""

    def __init__(self, contract: str = "blas", build: str = "blas"):
        for v in (contract, build):
            if not (v in ("blas", "pairwise") or (v.startswith("strided:") and int(v[8:]) >= 1)):
                raise ValueError(f"unknown summation order {v!r}")
        self.new = {"contract": contract, "build": build}

    def __enter__(self):
        self.old = dict(_ORDER)
        _ORDER.update(self.new)
        return self

    def __exit__(self, *exc):
        _ORDER.update(self.old)


def _ordered_matvec(A: np.ndarray, v: np.ndarray, order: str) -> np.ndarray:
    """A @ v summed over A's columns in the named order."""
    if order == "blas":
        return A @ v
    if order == "pairwise":
        return np.sum(A * v[None, :], axis=1)
    n = int(order[8:])
    s = A[:, 0::n] @ v[0::n]
    for i in range(1, n):
        s = s + A[:, i::n] @ v[i::n]
    return s


def build_P(prob: Problem, l: int, x) -> np.ndarray:
    """P = sum_i x_i A_i (schur.jl:36-39)."""
    k = prob.k[l]
    return _ordered_matvec(prob.Av[l], np.asarray(x, dtype=np.float64), _ORDER["build"]).reshape(k, k, order="F")


def contract(prob: Problem, l: int, M) -> np.ndarray:
    """[<A_i, M>]_i (schur.jl:41-44)."""
    return _ordered_matvec(prob.Av[l].T, np.asarray(M).ravel(order="F"), _ORDER["contract"])


def schur_build(prob: Problem, LX: list, MY: list, Y: list | None = None) -> np.ndarray:
    """Dense Schur S = sum_l Phat_l' Phat_l, Phat_i = L_X^-1 A_i M_Y (schur.jl:511-547).

    Blocks in ``prob.sparse`` take the sparse pair route (sparse_schur_block, reading Y;
    M_Y M_Y' when Y is not given).  Returns the full symmetric matrix; the reference keeps
    only the lower triangle authoritative (workspace.jl:1166-1172).
    """
    m = prob.m
    S = np.zeros((m, m))
    for l in range(prob.L):
        k = prob.k[l]
        if l in prob.sparse:
            ids, coo = prob.sparse[l]
            Yl = Y[l] if Y is not None else MY[l] @ MY[l].T
            scatter_sparse_schur(S, ids, sparse_schur_block(ids, coo, LX[l], Yl))
            continue
        MYl = np.tril(MY[l])
        Pm = np.empty((m, k * k))  # row i = vec(Phat_i)
        # constraints in chunks of ~256 MB: every column is computed as in one call, with less memory
        chunk = max(1, (1 << 25) // (k * k))
        for i0 in range(0, m, chunk):
            n = min(m, i0 + chunk) - i0
            panel = prob.Av[l][:, i0:i0 + n].T.reshape(n, k, k).transpose(0, 2, 1)  # panel[i] = A_i
            T = sla.solve_triangular(LX[l], np.concatenate(list(panel), axis=1), lower=True, check_finite=False)
            T = T.reshape(k, n, k).transpose(1, 0, 2)  # T[i] = L_X^-1 A_i
            Ph = T @ MYl                                # Phat_i = L_X^-1 A_i M_Y
            Pm[i0:i0 + n] = Ph.transpose(0, 2, 1).reshape(n, k * k)
        S += Pm @ Pm.T
    return S


# --------------------------------------------------------------------------
# Sparse blocks (types/constraints.jl:59-109, schur.jl:566-640, 936-1032, 1229-1318)
# --------------------------------------------------------------------------
@dataclass
class SparseBlockCOO:
    """SparseBlockCOO (types/constraints.jl:59-65): 1-based Int32 ptr / lin / row / col and val."""

    ptr: np.ndarray
    lin: np.ndarray
    row: np.ndarray
    col: np.ndarray
    val: np.ndarray


def block_active(Al: np.ndarray) -> np.ndarray:
    """Variables whose coefficient on the block is structurally nonzero, ascending: the
    reference's ``active`` / ``schur_order`` of a sparse block (test/sparse.jl:206-216)."""
    m = Al.shape[0]
    return np.flatnonzero(np.abs(np.asarray(Al).reshape(m, -1)).max(axis=1) > 0.0) if m else np.zeros(0, int)


def build_block_coo(Al: np.ndarray, order) -> SparseBlockCOO:
    """build_block_coo (types/constraints.jl:79-109): positions in ``order``, each
    coefficient's stored entries in column-major (CSC) order, both triangles, 1-based."""
    k = Al.shape[1]
    ptr, lin, row, col, val = [1], [], [], [], []
    for i in order:
        A = Al[i]
        for cc in range(k):
            for rr in range(k):
                if A[rr, cc] != 0.0:
                    row.append(rr + 1)
                    col.append(cc + 1)
                    lin.append(cc * k + rr + 1)
                    val.append(A[rr, cc])
        ptr.append(len(val) + 1)
    i32 = lambda v: np.asarray(v, dtype=np.int32)  # noqa: E731
    return SparseBlockCOO(i32(ptr), i32(lin), i32(row), i32(col), np.asarray(val, dtype=np.float64))


def ingest_sparse(c, A, C, blocks) -> Problem:
    """``ingest(c, A, C)`` with the listed blocks on the sparse route (ingest.jl sparse=true
    for those blocks): ids = block_active, coo = build_block_coo over ids."""
    prob = ingest(c, A, C)
    for l in blocks:
        Al = np.asarray(A[l], dtype=np.float64)
        ids = block_active(Al)
        prob.sparse[l] = (ids, build_block_coo(Al, ids))
    return prob


def two_sided_coo_product(left, coo: SparseBlockCOO, p: int, right) -> np.ndarray:
    """_two_sided_coo_product! (schur.jl:601-640): left * A_p * right as a sum of rank-one terms
    val_t left[:, row_t] right[col_t, :], first entry writing, later ones accumulating."""
    k = left.shape[0]
    t0, t1 = int(coo.ptr[p]) - 1, int(coo.ptr[p + 1]) - 1
    dest = np.zeros((k, k))
    for t in range(t0, t1):
        r, c = int(coo.row[t]) - 1, int(coo.col[t]) - 1
        term = np.outer(left[:, r], coo.val[t] * right[c, :])
        dest = term if t == t0 else dest + term
    return dest


def dot_dense_coo(M, coo: SparseBlockCOO, p: int) -> float:
    """_dot_dense_coo (schur.jl:566-579): sum of M[lin_t] val_t over p's entries, stored order."""
    flat = np.asarray(M).ravel(order="F")
    acc = 0.0
    for t in range(int(coo.ptr[p]) - 1, int(coo.ptr[p + 1]) - 1):
        acc += flat[int(coo.lin[t]) - 1] * coo.val[t]
    return acc


def sparse_schur_block(ids, coo: SparseBlockCOO, LX, Y) -> np.ndarray:
    """sparse_schur_block! (schur.jl:936-1032), general COO branch: W2 = X^-1 from L_X,
    W1_p = Y A_p W2, Svals packed over p <= r as <W1_p, A_r>."""
    k = LX.shape[0]
    W2 = kcholsolve(LX, np.eye(k))
    na = len(ids)
    out = []
    for p in range(na):
        W1 = two_sided_coo_product(Y, coo, p, W2)
        for r in range(p, na):
            out.append(dot_dense_coo(W1, coo, r))
    return np.asarray(out)


def scatter_sparse_schur(S, ids, Svals) -> None:
    """reduce_sparse_schur! scatter (schur.jl:1229-1318), lower-only form: S[ids[r], ids[p]] +=
    Svals over p <= r, mirrored here so S stays the full symmetric matrix."""
    na, q = len(ids), 0
    for p in range(na):
        for r in range(p, na):
            S[ids[r], ids[p]] += Svals[q]
            if r != p:
                S[ids[p], ids[r]] += Svals[q]
            q += 1


def schur_mul(S: np.ndarray, x) -> np.ndarray:
    """Symmetric S*x from the lower triangle (kkt.jl:3563-3577)."""
    Lw = np.tril(S)
    return Lw @ x + np.tril(S, -1).T @ x


# --------------------------------------------------------------------------
# KKT factor / solve / refinement (kkt.jl, kkt_backend.jl)
# --------------------------------------------------------------------------
# BLAS threads of the Schur Cholesky (None: leave the process setting).  The reference
# caps them at clamp(m / 256, 1, threads) around this phase (step.jl:124-157).
CHOLESKY_THREADS = None


def factor_dense_kkt(S: np.ndarray):
    """Schur Cholesky with the regularisation ladder (kkt.jl:1404-1464).

    Returns (L, ok, reg_attempts).
    """
    if CHOLESKY_THREADS is not None:
        from threadpoolctl import threadpool_limits

        with threadpool_limits(limits=int(CHOLESKY_THREADS)):
            return _factor_dense_kkt(S)
    return _factor_dense_kkt(S)


def _factor_dense_kkt(S: np.ndarray):
    m = S.shape[0]
    Sl = np.tril(S)
    c, info = lapack.dpotrf(np.asfortranarray(Sl), lower=1, clean=1)
    ok = info == 0
    attempts = 0
    reg = 0.0
    diag = np.abs(np.diag(S))
    while not ok and attempts < 6:
        attempts += 1
        reg = SQRT_EPS if attempts == 1 else reg * 10
        Sb = Sl.copy()
        Sb[np.arange(m), np.arange(m)] += reg * np.maximum(diag, 1.0)
        c, info = lapack.dpotrf(np.asfortranarray(Sb), lower=1, clean=1)
        ok = info == 0
    return (np.tril(c) if ok else None), ok, attempts


def solve_dense_kkt(Lf, r) -> np.ndarray:
    """dx = L^-T (L^-1 r) (kkt.jl:2570-2595, n_eq = 0); an ArrowFactor takes solve_arrow_kkt."""
    if isinstance(Lf, ArrowFactor):
        return solve_arrow_kkt(Lf, r)
    t = sla.solve_triangular(Lf, r, lower=True, check_finite=False)
    return sla.solve_triangular(Lf, t, lower=True, trans="T", check_finite=False)


# --------------------------------------------------------------------------
# Block-arrow elimination (workspace.jl:318-400, kkt.jl:1695-1716, 2101-2349, 2977-3096)
# --------------------------------------------------------------------------
def block_active_sets(prob: Problem) -> list:
    """Per block, the variables with a structurally nonzero coefficient, ascending (``cons.active``)."""
    out = []
    for l in range(prob.L):
        if l in prob.sparse:
            out.append(np.asarray(prob.sparse[l][0], dtype=np.int64))
        else:
            out.append(np.flatnonzero(np.any(prob.Av[l] != 0.0, axis=0)))
    return out


def arrow_partition(prob: Problem):
    """ArrowWorkspace (workspace.jl:318-400): a variable touching more than one block is global, one
    touching exactly one block is local to it.  None (the general route) when a variable touches no
    block or every variable is global.  Returns (global_ids, local_ids per block)."""
    m = prob.m
    freq = np.zeros(m, dtype=np.int64)
    owner = np.zeros(m, dtype=np.int64)
    for l, act in enumerate(block_active_sets(prob)):
        freq[act] += 1
        owner[act] = l
    if not np.all(freq > 0):
        return None
    gids = np.flatnonzero(freq > 1)
    if len(gids) >= m:
        return None
    lids = [np.flatnonzero((freq == 1) & (owner == l)) for l in range(prob.L)]
    return gids, lids


@dataclass
class ArrowFactor:
    gids: np.ndarray
    lids: list
    D: list        # lower Cholesky factor of each local block S[U_l, U_l]
    C: list        # coupling S[U_l, G]
    W: list        # D_l^-1 C_l
    Lred: np.ndarray  # lower factor of S[G, G] - sum_l C_l' W_l


def _chol_ladder(src: np.ndarray):
    """The relative-shift ladder of factor_arrow_kkt!'s local blocks (kkt.jl:2232-2260) and of
    _factor_with_relative_regularization! (kkt.jl:1695-1716): reg = sqrt(eps), x10, <= 6 attempts,
    diagonal += reg max(|src_ii|, 1).  Returns (L, ok, attempts)."""
    Lf, ok = kchol(src)
    attempts, reg = 0, SQRT_EPS
    n = src.shape[0]
    while not ok and attempts < 6:
        attempts += 1
        A = src.copy()
        A[np.arange(n), np.arange(n)] += reg * np.maximum(np.abs(np.diag(src)), 1.0)
        Lf, ok = kchol(A)
        reg *= 10
    return Lf, ok, attempts


def factor_arrow_kkt(S: np.ndarray, part):
    """factor_arrow_kkt! serial branch (kkt.jl:2232-2349): per local block chol(D_l) with its own
    ladder, W_l = D_l^-1 C_l (_prepare_arrow_coupling_solve!, :2101-2122), Sred = S_GG - sum_l
    C_l' W_l (_arrow_rank_sub_lower!), then the ladder on Sred.  Returns (factor, ok, attempts)."""
    gids, lids = part
    Ssym = np.tril(S) + np.tril(S, -1).T
    total = 0
    D, Cs, Ws = [], [], []
    Sred = Ssym[np.ix_(gids, gids)].copy()
    for ids in lids:
        if len(ids) == 0:
            D.append(None), Cs.append(None), Ws.append(None)
            continue
        Lf, ok, att = _chol_ladder(Ssym[np.ix_(ids, ids)])
        total += att
        if not ok:
            return None, False, total
        Cl = Ssym[np.ix_(ids, gids)]
        Wl = kcholsolve(Lf, Cl) if len(gids) else np.zeros((len(ids), 0))
        Sred -= Cl.T @ Wl
        D.append(Lf), Cs.append(Cl), Ws.append(Wl)
    Lred, ok, att = (np.zeros((0, 0)), True, 0) if len(gids) == 0 else _chol_ladder(np.tril(Sred) + np.tril(Sred, -1).T)
    total += att
    if not ok:
        return None, False, total
    return ArrowFactor(gids, lids, D, Cs, Ws, Lred), True, total


def solve_arrow_kkt(f: ArrowFactor, r) -> np.ndarray:
    """solve_arrow_kkt! serial branch (kkt.jl:2977-3096): r_G -= sum_l C_l' D_l^-1 r_U_l,
    x_G = Sred^-1 r_G, x_U_l = D_l^-1 r_U_l - W_l x_G."""
    r = np.asarray(r, dtype=np.float64)
    rg = r[f.gids].copy()
    tl = []
    for ids, Lf, Cl in zip(f.lids, f.D, f.C):
        if len(ids) == 0:
            tl.append(None)
            continue
        t = kcholsolve(Lf, r[ids])
        tl.append(t)
        rg -= Cl.T @ t
    out = np.zeros_like(r)
    if len(f.gids):
        xg = sla.solve_triangular(f.Lred, rg, lower=True, check_finite=False)
        xg = sla.solve_triangular(f.Lred, xg, lower=True, trans="T", check_finite=False)
        out[f.gids] = xg
    else:
        xg = np.zeros(0)
    for ids, t, Wl in zip(f.lids, tl, f.W):
        if len(ids):
            out[ids] = t - Wl @ xg
    return out


def factor_kkt(prob: Problem, S: np.ndarray, o: Options):
    """The Schur factor of the selected KKT route: dense Cholesky with the ladder, or the block-arrow
    elimination (an incompatible problem raises, as the reference's workspace does, workspace.jl:938-940)."""
    if o.kkt_route == "block_arrow":
        part = arrow_partition(prob)
        if part is None:
            raise ValueError("block_arrow route selected, but the problem is incompatible")
        return factor_arrow_kkt(S, part)
    return factor_dense_kkt(S)


def triangular_condition_estimate(L: np.ndarray) -> float:
    """_triangular_condition_estimate (mixed_precision_kkt.jl:371-392): LAPACK dtrcon of the lower
    factor in the one and infinity norms, cond = 1 / (rcond_1 rcond_inf) (inf when either is 0)."""
    a = np.asfortranarray(np.tril(L))
    r1, info1 = lapack.dtrcon(a, norm="1", uplo="L", diag="N")
    ri, info2 = lapack.dtrcon(a, norm="I", uplo="L", diag="N")
    if info1 != 0 or info2 != 0 or not (r1 > 0.0 and ri > 0.0):
        return math.inf
    est = 1.0 / (r1 * ri)
    return est if math.isfinite(est) else math.inf


def schur_condition_estimate(S: np.ndarray, L: np.ndarray) -> float:
    """cond_1(S) as LAPACK dpocon estimates it from the Cholesky factor: ||S||_1 of the symmetric S
    (from its lower triangle) over dpocon's reciprocal condition number."""
    Sl = np.tril(S)
    anorm = float(np.max(np.sum(np.abs(Sl + np.tril(Sl, -1).T), axis=0)))
    rc, info = lapack.dpocon(np.asfortranarray(np.tril(L)), anorm, uplo="L")
    if info != 0 or not rc > 0.0:
        return math.inf
    est = 1.0 / rc
    return est if math.isfinite(est) else math.inf


def equality_rank_tau(m, n, requested):
    """_equality_cholesky_rank_tau, Float64 branch (kkt.jl:103-112)."""
    req = requested if requested > 0 else SQRT_EPS
    req = min(SQRT_EPS, req)
    return max(max(m, n) * EPS, min(SQRT_EPS, req))


def factor_equality(Lf: np.ndarray, B: np.ndarray, requested: float):
    """n > 0 branch of _factor_dense_kkt_native! (kkt.jl:1466-1510) on the Schur factor Lf:
    B~ = L^-1 B, B^ = B~ D with D_jj = 1/||B~_j||_inf (kkt.jl:281-304), Q = B^'B^, its Cholesky and the
    numerical-rank gate (kkt.jl:122-176).  Returns (status 0 ok / 1 failed / 2 rank deficient, Bh, D, LQ, tau)."""
    m, n = B.shape
    Bt = sla.solve_triangular(Lf, B, lower=True, check_finite=False)
    D = np.ones(n)
    for j in range(n):
        norm = float(np.max(np.abs(Bt[:, j]))) if m else 0.0
        sc = 1.0 if norm == 0.0 else 1.0 / norm
        if not math.isfinite(sc):
            sc = 1.0
        D[j] = sc
        if sc != 1.0:
            Bt[:, j] *= sc
    Q = Bt.T @ Bt
    tau = equality_rank_tau(m, n, requested)
    c, info = lapack.dpotrf(np.asfortranarray(np.tril(Q)), lower=1, clean=1)
    if info != 0:
        return 1, Bt, D, None, tau
    dg = np.abs(np.diag(c))
    mn, mx = float(dg.min()), float(dg.max())
    if not (math.isfinite(mn) and math.isfinite(mx) and mx > 0):
        return 1, Bt, D, None, tau
    if not mn > tau * mx:
        return 2, Bt, D, None, tau
    if mn <= math.sqrt(tau) * mx:
        for a in range(n):
            for b2 in range(a + 1, n):
                if np.array_equal(Bt[:, a], Bt[:, b2]):
                    return 2, Bt, D, None, tau
    return 0, Bt, D, np.tril(c), tau


def solve_equality(Lf, Bh, D, LQ, r, p):
    """_solve_dense_kkt_owned! n > 0 branch (kkt.jl:2578-2587): the solution of
    S dx - B dy = r, B' dx = p."""
    rt = sla.solve_triangular(Lf, r, lower=True, check_finite=False)
    q = D * p - Bh.T @ rt
    t = sla.solve_triangular(LQ, q, lower=True, check_finite=False)
    dyh = sla.solve_triangular(LQ, t, lower=True, trans="T", check_finite=False)
    dx = sla.solve_triangular(Lf, rt + Bh @ dyh, lower=True, trans="T", check_finite=False)
    return dx, D * dyh


def direction_residual(S, r, dx):
    """rho_r = r - S dx; returns (||rho_r||_inf, rho_r) (kkt.jl:3689-3708)."""
    rho = r - schur_mul(S, dx)
    return knrm_inf(rho), rho


def refine_direction(S, Lf, r, dx, o: Options):
    """_refine_native_direction! (kkt.jl:3891-3963). Returns (dx, steps, residual)."""
    if o.refine_policy == "fixed":
        if o.refine_steps <= 0:
            return dx, 0, 0.0
        steps = 0
        residual, rho = direction_residual(S, r, dx)
        for _ in range(o.refine_steps):
            best = dx
            dx = dx + solve_dense_kkt(Lf, rho)
            corrected, rho_new = direction_residual(S, r, dx)
            if not math.isfinite(corrected) or corrected > residual:
                dx = best
                break
            steps += 1
            residual, rho = corrected, rho_new
        return dx, steps, residual
    cap = o.refine_max_steps
    if cap <= 0:
        return dx, 0, 0.0
    rel = o.refine_tol if o.refine_tol > 0 else REFINE_DEFAULT_TOL_ULPS * EPS
    abstol = rel * max(knrm_inf(r), 1.0)
    steps = 0
    residual, rho = direction_residual(S, r, dx)
    for _ in range(cap):
        if residual <= abstol:
            break
        best = dx
        dx = dx + solve_dense_kkt(Lf, rho)
        corrected, rho_new = direction_residual(S, r, dx)
        if not math.isfinite(corrected) or corrected > residual:
            dx = best
            break
        steps += 1
        small = corrected > residual * REFINE_MIN_DECREASE
        residual, rho = corrected, rho_new
        if small:
            break
    return dx, steps, residual


# --------------------------------------------------------------------------
# Newton step (step.jl, kernels/threaded.jl serial branches)
# --------------------------------------------------------------------------
@dataclass
class StepState:
    P: list = field(default_factory=list)
    R: list = field(default_factory=list)
    LX: list = field(default_factory=list)
    MY: list = field(default_factory=list)
    d: np.ndarray | None = None
    S: np.ndarray | None = None
    Lf: np.ndarray | None = None
    dx: np.ndarray | None = None
    dX: list = field(default_factory=list)
    dY: list = field(default_factory=list)
    rho: np.ndarray | None = None


def compute_residuals(prob: Problem, st: StepState, x, X, Y, mu, o: Options, factor: bool):
    """step.jl:25-72 and threaded.jl:423-505. Returns (p_res, d_res, blocks_ok)."""
    st.P = [build_P(prob, l, x) - X[l] - prob.C[l] for l in range(prob.L)]
    d = prob.c.copy()
    for l in range(prob.L):
        d -= contract(prob, l, Y[l])
    st.d = d
    p_res = max((knrm_inf(P) for P in st.P), default=0.0)
    d_res = knrm_inf(d)
    use_affine = o.parameter_strategy == "adaptive"
    st.R = []
    for l in range(prob.L):
        R = -(X[l] @ Y[l])
        if not use_affine:
            R[np.diag_indices(prob.k[l])] += mu[l]
        st.R.append(R)
    ok = True
    if factor:
        st.LX, st.MY = [], []
        for l in range(prob.L):
            lx, o1 = kchol(X[l])
            my, o2 = kchol(Y[l])
            ok = ok and o1 and o2
            st.LX.append(lx)
            st.MY.append(my)
    return p_res, d_res, ok


def rhs_v(prob: Problem, st: StepState, Y):
    """Z = X^-1 (P Y - R); v_i = <A_i, Z> (step.jl:95-106)."""
    v = np.zeros(prob.m)
    for l in range(prob.L):
        Z = kcholsolve(st.LX[l], st.P[l] @ Y[l] - st.R[l])
        v += contract(prob, l, Z)
    return v


def direction_blocks(prob: Problem, st: StepState, Y):
    """dX = sum dx_i A_i + P; dY = sym(X^-1 (R - dX Y)) (threaded.jl:617-627)."""
    st.dX, st.dY = [], []
    for l in range(prob.L):
        dX = build_P(prob, l, st.dx) + st.P[l]
        dY = symmetrize(kcholsolve(st.LX[l], st.R[l] - dX @ Y[l]))
        st.dX.append(dX)
        st.dY.append(dY)


def corrector_R_fixed(prob: Problem, st: StepState, X, Y, mu, beta):
    """threaded_corrector_rhs! serial branch (threaded.jl:839-854)."""
    for l in range(prob.L):
        k = prob.k[l]
        rl = float(np.sum((X[l] + st.dX[l]) * (Y[l] + st.dY[l]))) / mu[l] / k
        gl = max(rl ** 2 if rl < 1 else rl, beta)
        R = -(X[l] @ Y[l]) - st.dX[l] @ st.dY[l]
        R[np.diag_indices(k)] += gl * mu[l]
        st.R[l] = R


def corrector_R_mehrotra(prob: Problem, st: StepState, X, Y, sigma, mu, block_local):
    """threaded_mehrotra_corrector_rhs! serial branch (threaded.jl:943-1003)."""
    for l in range(prob.L):
        k = prob.k[l]
        target = sigma * float(np.sum(X[l] * Y[l])) / k if block_local else sigma * mu
        R = -(X[l] @ Y[l]) - st.dX[l] @ st.dY[l]
        R[np.diag_indices(k)] += target
        st.R[l] = R


def complementarity_diagnostics(prob, st, X, Y, tX, tY, detect_uniformity):
    """step.jl:373-544 (serial branches)."""
    comp = 0.0
    aff = 0.0
    uniform = detect_uniformity
    ref_val, ref_dim = None, 0
    for l in range(prob.L):
        value = float(np.sum(X[l] * Y[l]))
        comp += value
        if detect_uniformity:
            dim = prob.k[l]
            if dim > 0:
                if not math.isfinite(value):
                    uniform = False
                elif ref_val is None:
                    ref_val, ref_dim = value, dim
                elif not (value == ref_val if dim == ref_dim else value / dim == ref_val / ref_dim):
                    uniform = False
        aff += float(np.sum((X[l] + tX * st.dX[l]) * (Y[l] + tY * st.dY[l])))
    return comp, aff, uniform


def fraction_line_search(prob, X, Y, dX, dY, pf, df):
    """fraction_to_boundary_search! (step.jl:1081-1098)."""
    bx = min((fraction_to_boundary_bound(X[l], dX[l]) for l in range(prob.L)), default=1.0)
    by = min((fraction_to_boundary_bound(Y[l], dY[l]) for l in range(prob.L)), default=1.0)
    bx, by = min(1.0, bx), min(1.0, by)
    return (pf * bx if bx < 1 else 1.0), (df * by if by < 1 else 1.0)


def backtracking_line_search(prob, X, Y, dX, dY, gamma, min_step, ratio):
    """line_search! (step.jl:1025-1075)."""
    def side(M, dM):
        t = 1.0
        while True:
            ok = all(trial_has_cholesky_margin(M[l], t, dM[l], ratio) for l in range(prob.L))
            if ok or t < min_step:
                return t
            t *= gamma
    return side(X, dX), side(Y, dY)


def resolved_step_rule(prob: Problem, rule: str) -> str:
    """threaded.jl:1589-1600"""
    if rule == "auto":
        return "fraction_to_boundary" if all(k <= 2 for k in prob.k) else "backtrack"
    return rule


def newton_step(prob: Problem, o_iter: Options, o: Options, x, X, Y, mu, ctrl: Controller,
                iteration, gap_rel, scale_p, scale_d, trace=None):
    """newton_step! (step.jl:643-1006) for the dense Schur route, n_eq = 0."""
    st = StepState()
    p_res, d_res, ok = compute_residuals(prob, st, x, X, Y, mu, o_iter, factor=True)
    if not ok:
        return dict(status="breakdown", reason="Cholesky factorization of X or Y failed", p_res=p_res, d_res=d_res)
    st.S = schur_build(prob, st.LX, st.MY, Y)
    if trace is not None:
        trace.setdefault("S", []).append(st.S.copy())
        trace.setdefault("X", []).append([b.copy() for b in X])
        trace.setdefault("Y", []).append([b.copy() for b in Y])
    st.Lf, fok, reg_attempts = factor_kkt(prob, st.S, o)
    if not fok:
        return dict(status="breakdown", reason="Schur complement not positive definite", p_res=p_res, d_res=d_res,
                    reg_attempts=reg_attempts)
    # predictor
    v = rhs_v(prob, st, Y)
    r = -(st.d + v)
    st.dx = solve_dense_kkt(st.Lf, r)
    direction_blocks(prob, st, Y)
    adaptive = ctrl.strategy == "adaptive" and not ctrl.fallback
    kdim = max(sum(prob.k), 1)
    if adaptive:
        ap, ad = fraction_line_search(prob, X, Y, st.dX, st.dY, 1.0, 1.0)
        comp, aff, uniform = complementarity_diagnostics(prob, st, X, Y, ap, ad, True)
    else:
        ap, ad = 1.0, 1.0
        comp, aff, uniform = complementarity_diagnostics(prob, st, X, Y, 1.0, 1.0, False)
        uniform = False
    mu_g = comp / kdim
    mu_aff = max(0.0, aff / kdim)
    # controller diagnostics (step.jl:746-843)
    prev_p = _history_value(ctrl.history, "primal_step", 1.0)
    prev_d = _history_value(ctrl.history, "dual_step", 1.0)
    prev_bt = _history_value(ctrl.history, "backtracking_count", 0)
    prev_ref = _history_value(ctrl.history, "refinement_count", 0)
    cur_feas = max(p_res / max(scale_p, 1.0), d_res / max(scale_d, 1.0))
    if not ctrl.history:
        achieved = 1.0
    else:
        row = ctrl.history[-1]
        prev_feas = max(row["primal_residual"] / max(scale_p, 1.0), row["dual_residual"] / max(scale_d, 1.0), EPS)
        achieved = cur_feas / prev_feas
    pm = min((_diag_quality(L) for L in st.LX), default=1.0)
    dm = min((_diag_quality(L) for L in st.MY), default=1.0)
    quality = _diag_quality(st.Lf)
    regularization = 0.0 if reg_attempts <= 0 else SQRT_EPS * 10.0 ** (reg_attempts - 1)
    diag = Diagnostics(
        iteration=iteration, primal_residual=p_res / max(scale_p, 1.0), dual_residual=d_res / max(scale_d, 1.0),
        relative_gap=gap_rel, mu=mu_g, mu_aff=mu_aff, affine_primal_step=ap, affine_dual_step=ad,
        previous_primal_step=prev_p, previous_dual_step=prev_d, backtracking_count=prev_bt,
        regularization=regularization, refinement_count=prev_ref, factorization_quality=quality,
        achieved_residual_reduction=achieved, primal_psd_margin=pm, dual_psd_margin=dm,
    )
    params = ctrl.select(diag)
    # corrector
    if adaptive and not params.fallback:
        corrector_R_mehrotra(prob, st, X, Y, params.sigma, mu_g, block_local=not uniform)
    else:
        corrector_R_fixed(prob, st, X, Y, mu, o_iter.beta)
    v = rhs_v(prob, st, Y)
    r = -(st.d + v)
    st.dx = solve_dense_kkt(st.Lf, r)
    o_corr = replace(o_iter, refine_tol=params.refinement_tolerance,
                     refine_max_steps=params.refinement_max_count) if (adaptive and not params.fallback) else o_iter
    st.dx, refine_steps, _ = refine_direction(st.S, st.Lf, r, st.dx, o_corr)
    res, st.rho = direction_residual(st.S, r, st.dx)
    tol = max(SQRT_EPS, o.eps_primal, o.eps_dual) * max(knrm_inf(r), 1.0)
    if not math.isfinite(res) or res > tol:
        return dict(status="breakdown", reason="final structured KKT direction residual exceeded", p_res=p_res,
                    d_res=d_res, reg_attempts=reg_attempts)
    block_primal_residual = max((knrm_inf(P) for P in st.P), default=0.0)
    direction_blocks(prob, st, Y)
    return dict(status="ok", st=st, p_res=p_res, d_res=d_res, params=params, comp=comp, mu=mu_g, mu_aff=mu_aff,
                refine_steps=refine_steps, refine_residual=res, reg_attempts=reg_attempts,
                block_primal_residual=block_primal_residual)


def _diag_quality(Lm) -> float:
    """_cholesky_diagonal_quality (step.jl:233-245); an arrow factor is rated by its reduced factor
    alone (_kkt_factorization_quality, step.jl:266-300)."""
    if isinstance(Lm, ArrowFactor):
        Lm = Lm.Lred
    if Lm is None or Lm.size == 0:
        return 1.0
    dg = np.abs(np.diag(Lm))
    mn, mx = float(dg.min()), float(dg.max())
    return min(max(mn / mx, 0.0), 1.0) if (mx > 0 and math.isfinite(mx)) else 0.0


def solution_residuals(prob: Problem, x, X, Y):
    """validation.jl:15-58 (n_eq = 0)."""
    p_res = max((knrm_inf(build_P(prob, l, x) - X[l] - prob.C[l]) for l in range(prob.L)), default=0.0)
    d = prob.c.copy()
    for l in range(prob.L):
        d -= contract(prob, l, Y[l])
    return p_res, knrm_inf(d)


# --------------------------------------------------------------------------
# Setup: Ruiz equilibration (ingest.jl:1141-1270, 1432-1560) and the
# identity-metric cold start (interior_point.jl:699-1287, cold_start.jl)
# --------------------------------------------------------------------------
@dataclass
class Equilibration:
    """ingest.jl:1120-1139"""

    E: list
    s: np.ndarray
    objective_scale: float
    passes: list


def equilibrate(prob: Problem, ruiz_iters="auto"):
    """Dense equilibrate (ingest.jl:1432-1521) -> (scaled problem, Equilibration)."""
    if ruiz_iters == "auto":
        min_it, max_it, tol, adaptive = 2, 8, 1.0 / 20.0, True   # _ruiz_control, ingest.jl:1141-1160
    else:
        min_it = max_it = int(ruiz_iters)
        tol, adaptive = 0.0, False
    m = prob.m
    Av = [a.copy() for a in prob.Av]
    C2 = [c.copy() for c in prob.C]
    E, passes = [], []
    for l in range(prob.L):
        k = prob.k[l]
        El = np.ones(k)
        done = 0
        # T3[i][c, r] = A_i[r, c] (column i of the panel is vec(A_i), column-major)
        T3 = np.array(Av[l].T).reshape(m, k, k)
        for it in range(1, max_it + 1):
            done = it
            rn = np.max(np.abs(C2[l]), axis=1)                       # _ruiz_absorb_c_rowmax!
            if m:
                rn = np.maximum(rn, np.max(np.abs(T3), axis=(0, 1)))  # row r of every A_i
            converged = adaptive and it >= min_it and _ruiz_converged(rn, tol)
            e = np.array([1.0 / math.sqrt(v) if v > 0 else 1.0 for v in rn])   # _ruiz_step
            F = np.outer(e, e)                                           # F[r, c] = e[r] * e[c]
            C2[l] = C2[l] * F                                            # _ruiz_congruence!
            T3 *= F.T[None, :, :]                                        # A_i[r, c] *= e[r] * e[c]
            El = El * e
            if converged:
                break
        Av[l] = np.ascontiguousarray(T3.reshape(m, k * k).T)
        E.append(El)
        passes.append(done)
    maxima = np.zeros(m)
    for l in range(prob.L):
        if m:
            maxima = np.maximum(maxima, np.max(np.abs(Av[l]), axis=0))
    s = np.where(maxima > 0, maxima, 1.0)                             # _equilibration_variable_scales
    cc = prob.c / s
    osc = float(np.max(np.abs(cc))) if m else 0.0
    if not (osc > 0 and math.isfinite(osc)):
        osc = 1.0
    cc = cc / osc
    for l in range(prob.L):
        Av[l] /= s[None, :]
    return Problem(c=cc, C=C2, Av=Av, k=list(prob.k)), Equilibration(E, s, osc, passes)


def _ruiz_converged(rn, tol) -> bool:
    """ingest.jl:1162-1176"""
    dev = 0.0
    for v in rn:
        if not v > 0:
            continue
        if not math.isfinite(v):
            return False
        dev = max(dev, abs(math.log(v)))
    return dev <= tol


def unequilibrate(eq: Equilibration, x, X, Y):
    """ingest.jl:1523-1560 (n_eq = 0): x = x^ ./ s, X = E^-1 X^ E^-1, Y = E Y^ E objective_scale."""
    x = np.asarray(x) / eq.s
    Xo = [Xl / np.outer(E, E) for Xl, E in zip(X, eq.E)]
    Yo = [Yl * np.outer(E, E) for Yl, E in zip(Y, eq.E)]
    if eq.objective_scale != 1.0:
        Yo = [Yl * eq.objective_scale for Yl in Yo]
    return x, Xo, Yo


def cold_safety(scale):
    """cold_start.jl:24-26"""
    return SQRT_EPS * max(1.0, scale)


def cold_rounding_slack(scale):
    """cold_start.jl:47-49"""
    return 8.0 * EPS * max(1.0, scale)


def cold_psd_lower_bound(A) -> float:
    """cold_start.jl:141-181: exact for 1x1 and 2x2, symmetric Gershgorin on the lower triangle above."""
    n = A.shape[0]
    if n == 1:
        return float(A[0, 0])
    if n == 2:
        center = (A[0, 0] + A[1, 1]) / 2.0
        return float(center - math.hypot((A[0, 0] - A[1, 1]) / 2.0, A[1, 0]))
    bound = 0.0
    for row in range(n):
        radius = 0.0
        for col in range(n):
            if col != row:
                radius += abs(A[max(row, col), min(row, col)])
        cand = A[row, row] - radius
        bound = cand if row == 0 else min(bound, cand)
    return float(bound)


def cold_psd_stats(A):
    """cold_start.jl:184-201 -> (finite, scale, margin)"""
    n = A.shape[0]
    scale = 0.0
    for c in range(n):
        for r in range(c, n):
            v = A[r, c]
            if not math.isfinite(v):
                return False, scale, 0.0
            scale = max(scale, abs(v))
    return True, scale, cold_psd_lower_bound(A)


def cold_psd_shift(A):
    """_cold_start_psd_shift! (cold_start.jl:288-312), in place -> (ok, shift, margin, scale)."""
    n = A.shape[0]
    if A.ndim != 2 or A.shape[1] != n:
        raise ValueError("cold-start PSD shift requires a square matrix")
    if n == 0:
        return True, 0.0, 0.0, 1.0
    finite, scale, margin = cold_psd_stats(A)
    if not finite:
        return False, 0.0, margin, scale
    safety = cold_safety(scale)
    slack = cold_rounding_slack(scale)
    if margin > safety:
        return True, 0.0, margin, scale
    applied = 0.0
    for _ in range(16):
        shift = safety - margin + slack
        if not math.isfinite(shift):
            return False, applied, margin, scale
        A[np.diag_indices(n)] += shift
        applied += shift
        finite, scale, margin = cold_psd_stats(A)
        if not finite:
            return False, applied, margin, scale
        safety = cold_safety(scale)
        if margin > safety:
            return True, applied, margin, scale
    return False, applied, margin, scale


def cold_identity_mass_shifts(primal, dual, rho):
    """cold_start.jl:314-350 -> (ok, primal_shift, dual_shift)"""
    if not (rho > 0 and math.isfinite(primal) and math.isfinite(dual)):
        return False, 0.0, 0.0
    target = float(rho)
    p = max(0.0, (target - primal) / target)
    d = max(0.0, (target - dual) / target)
    ok = math.isfinite(p) and math.isfinite(d)
    return ok, (p if ok else 0.0), (d if ok else 0.0)


def cold_centering_shifts(kappa, primal, dual):
    """cold_start.jl:409-428 -> (ok, primal_shift, dual_shift)"""
    ok = (math.isfinite(kappa) and kappa > 0 and math.isfinite(primal) and primal > 0
          and math.isfinite(dual) and dual > 0)
    ps = kappa / (2.0 * dual) if ok else 0.0
    ds = kappa / (2.0 * primal) if ok else 0.0
    ok = ok and math.isfinite(ps) and math.isfinite(ds)
    return ok, (ps if ok else 0.0), (ds if ok else 0.0)


def cold_start(prob: Problem, o: Options):
    """_kkt_cold_start_initialization (interior_point.jl:699-1287), n_eq = 0.

    Returns (ok, reason, x, X, Y, mu, record)."""
    m, L, k = prob.m, prob.L, prob.k
    degree = sum(k)
    eye = [np.eye(kl) for kl in k]
    S = schur_build(prob, eye, eye)
    Lf, ok, reg = factor_kkt(prob, S, o)
    rec = {"regularization_attempts": reg}
    if not ok:
        return False, "identity_kkt_factorization", None, None, None, None, rec
    threshold = max(SQRT_EPS, o.eps_primal, o.eps_dual)
    refine_total = 0

    def solve_gated(r):
        nonlocal refine_total
        dx = solve_dense_kkt(Lf, r)
        res, _ = direction_residual(S, r, dx)
        scale = max(knrm_inf(r), 1.0)
        normalized = res / scale
        if math.isfinite(normalized) and normalized > threshold:
            dx, steps, _ = refine_direction(S, Lf, r, dx, replace(o, refine_policy="auto"))
            refine_total += steps
            res, _ = direction_residual(S, r, dx)
            normalized = res / scale
        return dx, normalized

    primal_rhs = np.zeros(m)
    for l in range(L):
        primal_rhs += contract(prob, l, prob.C[l])
    x, pn = solve_gated(primal_rhs)
    rec["primal_kkt_residual"] = pn
    if not (math.isfinite(pn) and pn <= threshold):
        return False, "kkt_cold_start_solve", None, None, None, None, rec
    v, dn = solve_gated(prob.c.copy())
    rec["dual_kkt_residual"] = dn
    rec["refinement_steps"] = refine_total
    if not (math.isfinite(dn) and dn <= threshold):
        return False, "kkt_cold_start_solve", None, None, None, None, rec
    X = [build_P(prob, l, x) - prob.C[l] for l in range(L)]
    Y = [build_P(prob, l, v) for l in range(L)]
    p_res, d_res = solution_residuals(prob, x, X, Y)
    scale_p = 1.0 + max((knrm_inf(Cl) for Cl in prob.C), default=0.0)
    scale_d = 1.0 + knrm_inf(prob.c)
    rec["primal_residual"], rec["dual_residual"] = p_res / scale_p, d_res / scale_d
    if not (math.isfinite(p_res / scale_p) and math.isfinite(d_res / scale_d)
            and p_res / scale_p <= threshold and d_res / scale_d <= threshold):
        return False, "kkt_cold_start_residuals", None, None, None, None, rec
    pshift, dshift = [], []
    for l in range(L):
        okp, ps, _, _ = cold_psd_shift(X[l])
        okd, ds, _, _ = cold_psd_shift(Y[l])
        if not (okp and okd):
            return False, "psd_shift", None, None, None, None, rec
        pshift.append(ps)
        dshift.append(ds)
    pmass = sum(float(np.trace(Xl)) for Xl in X)
    dmass = sum(float(np.trace(Yl)) for Yl in Y)
    okf, pf, df = cold_identity_mass_shifts(pmass, dmass, degree)
    if not okf:
        return False, "identity_mass", None, None, None, None, rec
    for l in range(L):
        if pf > 0:
            X[l][np.diag_indices(k[l])] += pf
        if df > 0:
            Y[l][np.diag_indices(k[l])] += df
    kappa_floor = sum(float(np.sum(X[l] * Y[l])) for l in range(L))
    pmass = sum(float(np.trace(Xl)) for Xl in X)
    dmass = sum(float(np.trace(Yl)) for Yl in Y)
    okc, pc, dc = cold_centering_shifts(kappa_floor, pmass, dmass)
    if not okc:
        return False, "centering", None, None, None, None, rec
    for l in range(L):
        if pc > 0:
            X[l][np.diag_indices(k[l])] += pc
        if dc > 0:
            Y[l][np.diag_indices(k[l])] += dc
    rec["primal_margin"] = min(cold_psd_stats(Xl)[2] for Xl in X)
    rec["dual_margin"] = min(cold_psd_stats(Yl)[2] for Yl in Y)
    rec["largest_primal_shift"] = max(pshift) + max(pf, 0.0) + max(pc, 0.0)
    rec["largest_dual_shift"] = max(dshift) + max(df, 0.0) + max(dc, 0.0)
    mu = [o.beta * float(np.sum(X[l] * Y[l])) / k[l] for l in range(L)]
    return True, "none", x, X, Y, mu, rec


def dual_objective(prob: Problem, Y) -> float:
    """interior_point.jl:7-14"""
    return float(sum(np.sum(prob.C[l] * Y[l]) for l in range(prob.L)))


@dataclass
class Result:
    status: str
    x: np.ndarray
    X: list
    Y: list
    pobj: float
    dobj: float
    iterations: int
    p_res: float
    d_res: float
    gap_rel: float
    trace: dict | None = None
    history: list | None = None


def solve(prob: Problem, o: Options, x0=None, X0=None, Y0=None, trace_schur=False, start: str = "given") -> Result:
    """_solve_sdp_core! warm-start path and loop (interior_point.jl:1323-2420), n_eq = 0, scaling none.

    start="cold": the identity-metric KKT cold start (interior_point.jl:699-1287) replaces x0/X0/Y0."""
    m, L, k = prob.m, prob.L, prob.k
    x = np.zeros(m) if x0 is None else np.asarray(x0, dtype=np.float64).copy()
    if start == "cold":
        ok, reason, x, X, Y, mu0, _ = cold_start(prob, o)
        if not ok:
            return Result("NumericalBreakdown", np.zeros(m), [np.eye(kl) for kl in k], [np.eye(kl) for kl in k],
                          0.0, 0.0, 0, math.inf, math.inf, math.inf, None, [])
    elif X0 is None:
        X = [o.omega_p * np.eye(kl) for kl in k]
        Y = [o.omega_d * np.eye(kl) for kl in k]
    else:
        X = [np.array(b, dtype=np.float64) for b in X0]
        Y = [np.array(b, dtype=np.float64) for b in Y0]
    mu = [o.beta * float(np.sum(X[l] * Y[l])) / k[l] for l in range(L)]
    ctrl = Controller(o)
    trace = {} if trace_schur else None
    pobj = float(prob.c @ x)
    dobj = dual_objective(prob, Y)
    st0 = StepState()
    p_res, d_res, _ = compute_residuals(prob, st0, x, X, Y, mu, o, factor=False)
    scale_p = 1.0 + max((knrm_inf(Cl) for Cl in prob.C), default=0.0)
    scale_d = 1.0 + knrm_inf(prob.c)
    stag = Stagnation(o.stall_iterations, o.stall_tolerance)
    best_merit = math.inf
    initial_merit = math.inf
    best = None
    comp_cur = float(sum(np.sum(X[l] * Y[l]) for l in range(L)))
    iteration = 0
    restarts = 0
    centering = 0
    status = "NotStarted"
    gap_rel = math.inf
    history = []
    while True:
        gap = pobj - dobj
        gap_rel = abs(gap) / max(1.0, (abs(pobj) + abs(dobj)) / 2)
        term_ok = p_res / scale_p <= o.eps_primal and d_res / scale_d <= o.eps_dual
        gap_ok = gap_rel <= o.eps_gap
        if term_ok and gap_ok:
            p_res, d_res = solution_residuals(prob, x, X, Y)
            term_ok = p_res / scale_p <= o.eps_primal and d_res / scale_d <= o.eps_dual
        merit = Stagnation.merit(o, p_res, d_res, gap_rel, comp_cur, scale_p, scale_d, max(abs(pobj), abs(dobj)))
        floor = Stagnation.at_precision_floor(p_res, d_res, gap_rel, scale_p, scale_d)
        stagnated = stag.observe(merit, floor, o.iter_max - iteration)
        if math.isfinite(merit) and not math.isfinite(initial_merit):
            initial_merit = merit
        if math.isfinite(merit) and merit < best_merit:
            best_merit = merit
            best = (x.copy(), [b.copy() for b in X], [b.copy() for b in Y], pobj, dobj, p_res, d_res, gap_rel)
        if term_ok and gap_ok:
            status = "Optimal"
            break
        if iteration >= o.iter_max:
            status = "IterLimit"
            break
        if stagnated:
            status = "InsufficientPrecision" if stag.reason == "precision_floor" else "Stalled"
            break
        o_iter = ctrl.iteration_options(o)
        res = newton_step(prob, o_iter, o, x, X, Y, mu, ctrl, iteration + 1, gap_rel, scale_p, scale_d, trace)
        if res["status"] == "breakdown":
            status = "NumericalBreakdown"
            break
        st = res["st"]
        p_res, d_res = res["p_res"], res["d_res"]
        params: IterationParameters = res["params"]
        rule = resolved_step_rule(prob, o.step_rule)
        if rule == "fraction_to_boundary":
            tX, tY = fraction_line_search(prob, X, Y, st.dX, st.dY, params.primal_fraction_to_boundary,
                                          params.dual_fraction_to_boundary)
        else:
            ratio = SQRT_EPS if o.parameter_strategy == "adaptive" else 0.0
            tX, tY = backtracking_line_search(prob, X, Y, st.dX, st.dY, params.backtracking_factor,
                                              params.minimum_step, ratio)
        bcount = (estimate_backtracking_count(tX, params.backtracking_factor, rule)
                  + estimate_backtracking_count(tY, params.backtracking_factor, rule))
        primal_feasible = p_res / scale_p <= o.eps_primal
        dual_feasible = d_res / scale_d <= o.eps_dual
        x_stuck = tX < params.minimum_step
        y_stuck = tY < params.minimum_step
        if x_stuck and primal_feasible and not y_stuck:
            tX, x_stuck = 0.0, False
        if y_stuck and dual_feasible and not x_stuck:
            tY, y_stuck = 0.0, False
        if x_stuck or y_stuck:
            # interior_point.jl:2010-2217
            near = 1000.0
            if (gap_rel < near * o.eps_gap and p_res / scale_p < near * o.eps_primal
                    and d_res / scale_d < near * o.eps_dual):
                status = "Stalled"
                break
            made_progress = (math.isfinite(initial_merit) and best_merit < initial_merit * 1e-3
                             and best_merit < NEAR_SOLUTION_MERIT)
            if o.stall_iterations > 0 and made_progress:
                status = "InsufficientPrecision"
                break
            if centering < o.max_centering and ctrl.beta < CENTERING_BETA_MAX:
                centering += 1
                ctrl.beta = min(ctrl.beta * CENTERING_BETA_STEP, CENTERING_BETA_MAX)
                mu = [ctrl.beta * float(np.sum(X[l] * Y[l])) / k[l] for l in range(L)]
                continue
            rescale_X = x_stuck and not primal_feasible
            rescale_Y = y_stuck and not dual_feasible
            if o.restart and restarts < o.max_restarts and (rescale_X or rescale_Y):
                restarts += 1
                if rescale_X:
                    X = [b * params.restart_scale for b in X]
                if rescale_Y:
                    Y = [b * params.restart_scale for b in Y]
                comp_cur = 0.0
                for l in range(L):
                    bc = float(np.sum(X[l] * Y[l]))
                    comp_cur += bc
                    mu[l] = ctrl.beta * bc / k[l]
                continue
            status = "Stalled" if (o.restart and restarts < o.max_restarts) else (
                "MaxRestartsExceeded" if restarts >= o.max_restarts else "NumericalBreakdown")
            break
        # update (threaded.jl:665-747)
        X = [X[l] + tX * st.dX[l] for l in range(L)]
        Y = [Y[l] + tY * st.dY[l] for l in range(L)]
        block_comp = [float(np.sum(X[l] * Y[l])) for l in range(L)]
        comp_after = float(sum(block_comp))
        finite = all(np.all(np.isfinite(b)) for b in X + Y)
        x = x + tX * st.dx
        if not (finite and np.all(np.isfinite(x)) and all(math.isfinite(v) for v in mu)):
            status = "NumericalBreakdown"
            break
        # accepted residuals (interior_point.jl:1287-1311)
        p_after = abs(1.0 - tX) * res["block_primal_residual"]
        d_vec = (1.0 - tY) * st.d - tY * st.rho
        d_after = knrm_inf(d_vec)
        ctrl.record(iteration=iteration + 1, complementarity_before=res["comp"], complementarity_after=comp_after,
                    primal_residual=p_res, dual_residual=d_res, primal_step=tX, dual_step=tY,
                    primal_residual_after=p_after, dual_residual_after=d_after, backtracking_count=bcount,
                    refinement_count=res["refine_steps"], selected=params)
        comp_cur = comp_after
        adaptive_mu = ctrl.strategy == "adaptive" and not ctrl.fallback
        if adaptive_mu:
            g = ctrl.beta * comp_after / max(sum(k), 1)
            mu = [g for _ in range(L)]
        else:
            mu = [ctrl.beta * block_comp[l] / k[l] for l in range(L)]
        pobj = float(prob.c @ x)
        dobj = dual_objective(prob, Y)
        iteration += 1
        history.append(dict(iteration=iteration, pobj=pobj, dobj=dobj, p_res=res["p_res"], d_res=res["d_res"],
                            tX=tX, tY=tY, sigma=params.sigma, mu=res["mu"], mu_aff=res["mu_aff"],
                            reg_attempts=res["reg_attempts"], refine_steps=res["refine_steps"],
                            fallback=ctrl.fallback))
        p_res, d_res = p_after, d_after
    if status != "Optimal" and best is not None:
        x, X, Y, pobj, dobj, p_res, d_res, gap_rel = best
    return Result(status, x, X, Y, pobj, dobj, iteration, p_res, d_res, gap_rel, trace, history)


def solve_equilibrated(prob: Problem, o: Options, ruiz_iters="auto", start: str = "cold", x0=None, X0=None,
                       Y0=None) -> tuple:
    """equilibrate -> solve -> unequilibrate (the reference's :sdp_ruiz scaling policy around
    _solve_sdp_core!); objectives are recomputed on the original problem.  Returns (Result, Equilibration)."""
    sp, eq = equilibrate(prob, ruiz_iters)
    if start != "cold" and X0 is not None:
        # a warm start mapped into the scaled space (interior_point.jl:606-661)
        x0 = np.asarray(x0) * eq.s
        X0 = [Xl * np.outer(E, E) for Xl, E in zip(X0, eq.E)]
        Y0 = [Yl / np.outer(E, E) / eq.objective_scale for Yl, E in zip(Y0, eq.E)]
    r = solve(sp, o, x0=x0, X0=X0, Y0=Y0, start=start)
    x, X, Y = unequilibrate(eq, r.x, r.X, r.Y)
    pobj = float(prob.c @ x)
    dobj = dual_objective(prob, Y)
    gap = abs(pobj - dobj) / max(1.0, (abs(pobj) + abs(dobj)) / 2)
    p_res, d_res = solution_residuals(prob, x, X, Y)
    return Result(r.status, x, X, Y, pobj, dobj, r.iterations, p_res, d_res, gap, None, r.history), eq


def certificate(prob: Problem, result: Result, tol_primal, tol_dual, tol_gap):
    """Original-coordinate certificate for the dual-form model (public/optimize.jl:520-662).

    The user's model is max b'y s.t. Z = C - sum y_i A_i >= 0 (core x = y,
    A_core = -A, C_core = -C, c = -b); the core slack X is the constraint value
    and the core Y is the constraint dual.
    """
    x, X, Y = result.x, result.X, result.Y
    p = 0.0
    d = 0.0
    rows = []
    for l in range(prob.L):
        Z = build_P(prob, l, x) - prob.C[l]
        rows.append(Z)
        p = max(p, max(0.0, -float(np.linalg.eigvalsh(Z).min())))
        d = max(d, max(0.0, -float(np.linalg.eigvalsh(Y[l]).min())))
    stationarity = prob.c.copy()
    for l in range(prob.L):
        stationarity -= contract(prob, l, Y[l])
    d = max(d, knrm_inf(stationarity))
    gap = abs(result.pobj - result.dobj) / max(1.0, (abs(result.pobj) + abs(result.dobj)) / 2)
    pscale = max(1.0, knrm_inf(x), max((knrm_inf(z) for z in rows), default=0.0))
    dscale = max(1.0, knrm_inf(prob.c), max((knrm_inf(b) for b in Y), default=0.0))
    finite = all(math.isfinite(v) for v in (p, d, gap, result.pobj, result.dobj))
    valid = (result.status == "Optimal" and finite and p / pscale <= tol_primal and d / dscale <= tol_dual
             and gap <= tol_gap)
    return dict(valid=valid, primal_residual=p / pscale, dual_residual=d / dscale, relative_gap=gap)
