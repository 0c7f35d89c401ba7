"""Host-side interior-point control: options, parameter policies, stagnation.

This is the scalar logic that stays on the host while every array phase runs
on the GPU.  Names and semantics mirror SDPX.jl:

* ``SolverOptions``            -- types/workspaces.jl:101-228 (defaults)
* ``FixedParameterPolicy``     -- adaptive_parameters.jl:21-32, 112-131, 160-186
* ``AdaptiveParameterPolicy``  -- adaptive_parameters.jl:40-53, 133-158, 253-453
* ``AdaptiveIPMController``    -- adaptive_parameters.jl:440-476, 510-548, 597-782
* ``StagnationDetector``       -- stagnation.jl:56-177
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

EPS = 2.220446049250313e-16
SQRT_EPS = math.sqrt(EPS)
REFINE_MIN_DECREASE = 0.5          # kkt.jl:3634
REFINE_DEFAULT_TOL_ULPS = 64       # kkt.jl:3637
NEAR_SOLUTION_MERIT = 100.0        # interior_point.jl:2703
CENTERING_BETA_STEP = 4.0          # interior_point.jl:2706
CENTERING_BETA_MAX = 0.5           # interior_point.jl:2709
PRECISION_FLOOR_ULPS = 16          # stagnation.jl:13
PROJECTION_SLACK = 10.0            # stagnation.jl:26


@dataclass(frozen=True)
class SolverOptions:
    """Float64 subset of SDPX's SolverOptions used by the dense SDP route."""

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
    refine_steps: int = 1
    refine_policy: str = "auto"
    refine_max_steps: int = 8
    refine_tol: float = 0.0
    parameter_strategy: str = "adaptive"
    adaptive_sigma_max: float = 0.0
    verbosity: int = 0
    # mixed precision (the FP32 -> FP64 instance of kernels/mixed_precision_kkt.jl): "fp64",
    # "mixed" or "mixed_bf16x3".  Mixed assembles the Schur Gram product in FP32 ("mixed") or as
    # three bf16 MFMA products of a hi / lo split of Phat ("mixed_bf16x3", about 16 significant
    # bits) while the factor's condition estimate stays below mixed_cond_limit (FP32) or
    # bf16x3_cond_limit, and refines every direction against the matrix-free FP64 Schur operator
    # to mixed_refine_target * max(||r||_inf, 1) in at most mixed_max_refine steps; otherwise the
    # iteration is redone in FP64 and FP64 is kept from then on.
    precision: str = "fp64"
    # "auto": the static condition cutoff (the xPOCON estimate of cond_1(S) against
    # mixed_cond_limit / bf16x3_cond_limit) before a reduced-precision step; "on": the reference's
    # measured trial (mixed_precision_kkt.jl:1008-1017, mode :on): no estimate, the predictor
    # residual guard and the monotone refinement within its budget decide
    mixed_mode: str = "auto"
    # factor the reduced-precision Schur matrix in FP32 too (FP32 MFMA panels and trailing
    # updates), under the same guards: the condition estimate of the FP32 factor against the
    # limit ("auto") or the refinement guards ("on"); a failing step is redone in FP64
    mixed_factor: bool = False
    # form P-hat on FP32 MFMA for the BF16x3 assembly (FP32 copy of the constraint panel, bf16
    # planes written by the product) instead of in FP64 followed by the split
    mixed_panels32: bool = True
    mixed_cond_limit: float = 1e6
    bf16x3_cond_limit: float = 1e5
    mixed_refine_target: float = 1e-10
    mixed_max_refine: int = 32
    # the batched route's refinement budget per direction and member: one refinement round is a
    # launch over every member still refining, so a slow member holds the whole batch; a member
    # that would need more rounds redoes its step in FP64 (2 measured fastest on the 4096-member
    # sparse and the 1024-member layer batches among 2, 3, 4, 6, 8, 12 and 32)
    mixed_batch_max_refine: int = 2
    # KKT route: "dense_cholesky" or "block_arrow" (the exact block-arrow elimination of
    # factor_arrow_kkt! / solve_arrow_kkt!, kkt.jl:2132-2349, 2977-3096; an incompatible problem
    # raises, as the reference's workspace does, workspace.jl:938-940)
    kkt_route: str = "dense_cholesky"

    def __post_init__(self):
        if not 0 < self.beta < 1:
            raise ValueError("beta must be strictly between zero and one")
        if not 0 < self.gamma < 1:
            raise ValueError("gamma must be strictly between zero and one")
        if self.parameter_strategy not in ("fixed", "adaptive"):
            raise ValueError("parameter_strategy must be 'fixed' or 'adaptive'")
        if self.step_rule not in ("auto", "backtrack", "fraction_to_boundary"):
            raise ValueError("step_rule must be 'backtrack', 'fraction_to_boundary', or 'auto'")
        if self.refine_policy not in ("fixed", "adaptive", "auto"):
            raise ValueError("refine_policy must be 'fixed', 'adaptive', or 'auto'")
        if self.kkt_route not in ("dense_cholesky", "block_arrow"):
            raise ValueError("kkt_route must be 'dense_cholesky' or 'block_arrow'")
        if self.precision not in ("fp64", "mixed", "mixed_bf16x3"):
            raise ValueError("precision must be 'fp64', 'mixed' or 'mixed_bf16x3'")
        if self.mixed_mode not in ("auto", "on"):
            raise ValueError("mixed_mode must be 'auto' or 'on'")


def settings(tolerance: float = 1e-8, **kw) -> SolverOptions:
    """Public-settings shortcut: one tolerance for gap, primal and dual (public/settings.jl:308-338)."""
    return SolverOptions(eps_gap=tolerance, eps_primal=tolerance, eps_dual=tolerance, **kw)


@dataclass
class IterationParameters:
    sigma: float
    primal_fraction_to_boundary: float
    dual_fraction_to_boundary: float
    backtracking_factor: float
    refinement_tolerance: float
    refinement_max_count: int
    minimum_step: float
    restart_scale: float
    regularization: float = 0.0
    fallback: bool = False
    fallback_reason: str = "none"


@dataclass
class FixedParameterPolicy:
    sigma: float
    primal_fraction: float
    dual_fraction: float
    backtracking: float
    refinement_tolerance: float
    refinement_max_count: int
    minimum_step: float
    restart_scale: float

    @classmethod
    def from_options(cls, o: SolverOptions) -> "FixedParameterPolicy":
        tol = o.refine_tol if o.refine_tol > 0 else REFINE_DEFAULT_TOL_ULPS * EPS
        count = o.refine_steps if o.refine_policy == "fixed" else o.refine_max_steps
        return cls(o.beta, o.gamma, o.gamma, o.gamma, tol, count, o.min_step, o.omega_step)

    def parameters(self, fallback: bool = False, reason: str = "none") -> IterationParameters:
        return IterationParameters(self.sigma, self.primal_fraction, self.dual_fraction, self.backtracking,
                                   self.refinement_tolerance, self.refinement_max_count, self.minimum_step,
                                   self.restart_scale, 0.0, fallback, reason)


@dataclass
class AdaptiveParameterPolicy:
    fallback: FixedParameterPolicy
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
    instability_limit: int = 2

    @classmethod
    def from_options(cls, o: SolverOptions) -> "AdaptiveParameterPolicy":
        fx = FixedParameterPolicy.from_options(o)
        req = min(o.eps_gap, o.eps_primal, o.eps_dual)
        floor = max(REFINE_DEFAULT_TOL_ULPS * EPS, req * req if req > 0 else 0.0)
        smax = max(o.adaptive_sigma_max, fx.sigma) if o.adaptive_sigma_max > 0 else max(0.5, fx.sigma)
        return cls(fx, min(1 / 50, fx.sigma), smax, min(4 / 5, fx.primal_fraction), max(99 / 100, fx.primal_fraction),
                   min(11 / 20, fx.backtracking), max(9 / 10, fx.backtracking), 2 / 25, 1 / 5, floor, req)


@dataclass
class IterationDiagnostics:
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

    def finite(self) -> bool:
        vals = (self.primal_residual, self.dual_residual, self.mu, self.mu_aff, self.affine_primal_step,
                self.affine_dual_step, self.factorization_quality, self.primal_psd_margin, self.dual_psd_margin)
        return all(math.isfinite(v) for v in vals) and self.mu > 0 and self.mu_aff >= 0


def _clamp(v, lo, hi):
    return min(max(v, lo), hi)


def _adaptive_fraction(p: AdaptiveParameterPolicy, affine_step, previous_step, backtracking_count, ratio):
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
    if ratio < 1e-6:
        cand = min(cand, 9 / 10)
    return _clamp(cand, p.fraction_min, p.fraction_max)


def select_parameters(policy, diag: IterationDiagnostics, history: list) -> IterationParameters:
    """Pure parameter selection (adaptive_parameters.jl:188-194, 253-453)."""
    if isinstance(policy, FixedParameterPolicy):
        return policy.parameters()
    p = policy
    if not diag.finite():
        return p.fallback.parameters(True, "nonfinite_diagnostics")
    if diag.factorization_quality <= 0:
        return p.fallback.parameters(True, "degraded_factorization")
    unstable = 0
    for row in reversed(history):
        if not row.get("unstable", False):
            break
        unstable += 1
    if unstable >= p.instability_limit:
        return p.fallback.parameters(True, "repeated_instability")
    if max(diag.primal_residual, diag.dual_residual, diag.relative_gap) >= 1:
        return p.fallback.parameters()
    ratio = _clamp(diag.mu_aff / diag.mu, 0.0, 1.0)
    amin = min(diag.affine_primal_step, diag.affine_dual_step)
    sigma_c = ratio ** (3 if amin >= 3 / 4 else 2)
    if amin < 1 / 20 or diag.achieved_residual_reduction > 5 / 4 or diag.factorization_quality < SQRT_EPS:
        floor = p.severe_centrality_floor
    elif (amin < 1 / 4 or diag.backtracking_count >= 4
          or diag.refinement_count >= max(p.fallback.refinement_max_count - 1, 1)):
        floor = p.centrality_floor
    else:
        floor = 0.0
    sigma_c = max(sigma_c, floor)
    if diag.iteration <= 1:
        sigma_c = max(sigma_c, min(p.fallback.sigma, p.sigma_max))
    prev_sigma = history[-1].get("sigma", p.fallback.sigma) if history else p.fallback.sigma
    sigma = _clamp(4 / 5 * sigma_c + 1 / 5 * prev_sigma, p.sigma_min, p.sigma_max)
    pf = _adaptive_fraction(p, diag.affine_primal_step, diag.previous_primal_step, diag.backtracking_count, ratio)
    df = _adaptive_fraction(p, diag.affine_dual_step, diag.previous_dual_step, diag.backtracking_count, ratio)
    if diag.backtracking_count >= 6 or diag.achieved_residual_reduction > 5 / 4:
        bt = 13 / 20
    elif diag.backtracking_count >= 3:
        bt = 18 / 25
    elif amin > 4 / 5:
        bt = 17 / 20
    else:
        bt = 4 / 5
    bt = _clamp(bt, p.backtracking_min, p.backtracking_max)
    reg = max(diag.regularization / 10, SQRT_EPS) if diag.regularization > 0 else 0.0
    rtol = max(p.refinement_floor, p.fallback.refinement_tolerance)
    degraded = diag.factorization_quality < SQRT_EPS
    rmax = p.fallback.refinement_max_count if (degraded or diag.refinement_count > 0) else min(
        p.fallback.refinement_max_count, 2)
    req = p.requested_tolerance if p.requested_tolerance > 0 else SQRT_EPS
    floor_step = max(100 * EPS, min(p.fallback.minimum_step, req / 100))
    min_step = p.fallback.minimum_step if (degraded or diag.backtracking_count >= 8) else floor_step
    fmin = max(min(diag.primal_residual, diag.dual_residual), req)
    fratio = max(diag.primal_residual, diag.dual_residual) / fmin
    rs = 1000.0 if fratio > 1e6 else (100.0 if fratio > 1000 else 10.0)
    rs = _clamp(rs, 10.0, p.fallback.restart_scale)
    return IterationParameters(sigma, pf, df, bt, rtol, rmax, min_step, rs, reg)


def outcome_is_unstable(iteration, reduction, fratio, primal_step, dual_step):
    """adaptive_parameters.jl:550-586"""
    if not all(math.isfinite(v) for v in (reduction, fratio, primal_step, dual_step)):
        return True, "nonfinite_outcome"
    if iteration > 3 and reduction > 2:
        return True, "complementarity_growth"
    if iteration > 3 and fratio > 3:
        return True, "feasibility_growth"
    if iteration > 3 and primal_step < 1e-6 and dual_step < 1e-6 and fratio >= 1:
        return True, "repeated_tiny_step"
    if iteration > 5 and min(primal_step, dual_step) < 1 / 50 and fratio > 49 / 50 and reduction > 49 / 50:
        return True, "stalled_progress"
    return False, "none"


@dataclass
class AdaptiveIPMController:
    strategy: str
    default_beta: float
    default_gamma: float
    beta: float
    gamma: float
    policy: object
    parameters: IterationParameters
    previous_feasibility: float = math.nan
    instability_count: int = 0
    fallback: bool = False
    history: list = field(default_factory=list)

    @classmethod
    def from_options(cls, o: SolverOptions) -> "AdaptiveIPMController":
        policy = (AdaptiveParameterPolicy.from_options(o) if o.parameter_strategy == "adaptive"
                  else FixedParameterPolicy.from_options(o))
        return cls(o.parameter_strategy, o.beta, o.gamma, o.beta, o.gamma, policy,
                   FixedParameterPolicy.from_options(o).parameters())

    def last(self, name, default):
        return self.history[-1].get(name, default) if self.history else default

    def select_iteration_parameters(self, diag: IterationDiagnostics) -> IterationParameters:
        if self.fallback:
            fx = FixedParameterPolicy(self.default_beta, self.default_gamma, self.default_gamma, self.default_gamma,
                                      self.parameters.refinement_tolerance, self.parameters.refinement_max_count,
                                      self.parameters.minimum_step, self.parameters.restart_scale)
            params = fx.parameters(True, "previous_instability")
        else:
            params = select_parameters(self.policy, diag, self.history)
        self.parameters = params
        self.beta = params.sigma
        self.gamma = min(params.primal_fraction_to_boundary, params.dual_fraction_to_boundary)
        if params.fallback:
            self.fallback = True
            self.beta = self.default_beta
            self.gamma = self.default_gamma
        return params

    def record_and_update(self, *, iteration, complementarity_before, complementarity_after, primal_residual,
                          dual_residual, primal_step, dual_step, primal_residual_after, dual_residual_after,
                          backtracking_count, refinement_count, selected: IterationParameters):
        reduction = complementarity_after / complementarity_before if complementarity_before > 0 else 1.0
        feas = max(primal_residual, dual_residual)
        prev = self.previous_feasibility
        fratio = feas / prev if (math.isfinite(prev) and prev > 0) else 1.0
        unstable, reason = outcome_is_unstable(iteration, reduction, fratio, primal_step, dual_step)
        if self.strategy == "adaptive" and not self.fallback:
            if reason in ("nonfinite_outcome", "complementarity_growth", "feasibility_growth"):
                self.instability_count = 2
            elif unstable:
                self.instability_count += 1
            else:
                self.instability_count = max(self.instability_count - 1, 0)
            if self.instability_count >= 2:
                self.beta, self.gamma, self.fallback = self.default_beta, self.default_gamma, True
        before = max(primal_residual, dual_residual, EPS)
        self.history.append(dict(
            iteration=iteration, sigma=selected.sigma, primal_step=primal_step, dual_step=dual_step,
            backtracking_count=backtracking_count, refinement_count=refinement_count,
            primal_residual=primal_residual, dual_residual=dual_residual,
            achieved_residual_reduction=max(primal_residual_after, dual_residual_after) / before,
            unstable=unstable, fallback=self.fallback or selected.fallback,
        ))
        self.previous_feasibility = feas

    def iteration_options(self, o: SolverOptions) -> SolverOptions:
        """controller_options (adaptive_parameters.jl:771-782)."""
        return replace(o, beta=self.beta, gamma=self.gamma,
                       parameter_strategy="fixed" if self.fallback else self.strategy)


def estimate_backtracking_count(step, gamma, rule) -> int:
    """adaptive_parameters.jl:500-508"""
    if rule != "backtrack" or not math.isfinite(step) or step <= 0 or step >= 1 or gamma <= 0 or gamma >= 1:
        return 0
    return max(0, int(round(math.log(step) / math.log(gamma))))


class StagnationDetector:
    """stagnation.jl:56-177"""

    def __init__(self, window: int, tolerance: float):
        self.window = max(window, 0)
        self.tolerance = max(tolerance, 0.0)
        self.merits: list[float] = []
        self.at_floor: list[bool] = []
        self.reason = "none"

    @staticmethod
    def merit(o: SolverOptions, p_res, d_res, gap_rel, comp, scale_p, scale_d, objective_scale) -> float:
        def safe(e):
            return e if e > 0 else EPS
        m = max((p_res / scale_p) / safe(o.eps_primal), (d_res / scale_d) / safe(o.eps_dual), gap_rel / safe(o.eps_gap))
        if math.isfinite(comp) and comp > 0:
            m = max(m, comp / (safe(o.eps_gap) * max(1.0, objective_scale)))
        return m

    @staticmethod
    def at_precision_floor(p_res, d_res, gap_rel, scale_p, scale_d) -> bool:
        f = PRECISION_FLOOR_ULPS * EPS
        return p_res / scale_p <= f and d_res / scale_d <= f and gap_rel <= f

    def observe(self, merit: float, floor: bool, remaining: int) -> bool:
        if self.window <= 0:
            return False
        self.merits.append(merit)
        self.at_floor.append(floor)
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
        if math.log(merit) / rate > PROJECTION_SLACK * max(remaining, 1):
            self.reason = "precision_floor" if all(self.at_floor) else "too_slow"
            return True
        return False
