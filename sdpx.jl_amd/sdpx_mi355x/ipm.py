"""Primal-dual interior-point driver of the dense SDP route on the MI355X backend.

Host control follows SDPX.jl's ``_solve_sdp_core!`` warm-start path and loop
(solver/interior_point.jl:1323-2420) and ``newton_step!`` (step.jl:643-1006);
every array phase is one call into the HIP library through
``MI355XDenseBackend``.  n_eq = 0, ``scaling = :none``.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from .backend import MI355XDenseBackend, SDPProblem
from .policy import (CENTERING_BETA_MAX, CENTERING_BETA_STEP, EPS, NEAR_SOLUTION_MERIT, REFINE_DEFAULT_TOL_ULPS,
                     SQRT_EPS, AdaptiveIPMController, IterationDiagnostics, SolverOptions, StagnationDetector,
                     estimate_backtracking_count)

PHASES = ("residual_and_block_factor", "schur_assembly", "kkt_factorization", "predictor", "complementarity_analysis",
          "corrector", "line_search", "update")


@dataclass
class SDPResult:
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
    message: str = ""
    phase_times: dict = field(default_factory=dict)
    history: list = field(default_factory=list)
    schur_trace: list = field(default_factory=list)  # first trace_schur steps: S and the iterate it was built at
    fp32_iterations: int = 0   # Newton steps whose Schur matrix was assembled in FP32 or BF16x3 (mixed precision)
    fp64_iterations: int = 0
    equilibration: object = None  # the Ruiz scaling the solve ran under (equilibrate=True)


def resolved_step_rule(k, rule: str) -> str:
    """threaded.jl:1589-1600"""
    if rule == "auto":
        return "fraction_to_boundary" if all(kl <= 2 for kl in k) else "backtrack"
    return rule


class _Clock:
    def __init__(self, backend, enabled):
        self.b = backend
        self.enabled = enabled
        self.t = {p: 0.0 for p in PHASES}
        self._last = time.perf_counter()

    def start(self):
        if self.enabled:
            self.b.synchronize()
        self._last = time.perf_counter()

    def lap(self, phase):
        if self.enabled:
            self.b.synchronize()
        now = time.perf_counter()
        self.t[phase] += now - self._last
        self._last = now


def newton_step(b: MI355XDenseBackend, o_iter: SolverOptions, o: SolverOptions, mu, ctrl: AdaptiveIPMController,
                iteration: int, gap_rel: float, scale_p: float, scale_d: float, clock: _Clock, trace_schur=None,
                use32: bool = False):
    """newton_step! (step.jl:643-1006) on the device backend.

    use32 (mixed precision): the Schur Gram product is assembled in FP32 and both directions are
    refined against the matrix-free FP64 Schur operator.  Returns status "retry_fp64" when the
    factor's condition estimate exceeds o.mixed_cond_limit or a direction misses its refinement
    target (the caller redoes the step in FP64)."""
    k = b.k
    clock.start()
    p_res, d_res, ok = b.compute_residuals(mu, affine=o_iter.parameter_strategy == "adaptive", factor=True)
    clock.lap("residual_and_block_factor")
    if not ok:
        return dict(status="breakdown", reason="Cholesky factorization of X or Y failed (iterate left the interior)",
                    p_res=p_res, d_res=d_res)
    if o.precision != "fp64":
        b.set_precision((1 if o.precision == "mixed" else 2) if use32 else 0)
        if o.mixed_factor:
            b.set_factor_precision(1 if use32 else 0)
        if o.precision == "mixed_bf16x3":
            b.set_panel_precision(1 if o.mixed_panels32 else 0)
    b.schur_build()
    clock.lap("schur_assembly")
    if trace_schur is not None:
        x_it, X_it, Y_it = b.get_iterate()
        trace_schur.append(dict(S=b.get_schur(), x=x_it, X=X_it, Y=Y_it))
    fok, reg_attempts = b.factorize()
    clock.lap("kkt_factorization")
    if not fok:
        if use32:
            return dict(status="retry_fp64", reason="reduced-precision Schur matrix not positive definite")
        return dict(status="breakdown",
                    reason=f"Schur complement not positive definite after {reg_attempts} regularization attempt(s)",
                    p_res=p_res, d_res=d_res, reg_attempts=reg_attempts)
    if use32 and o.mixed_mode == "auto":
        # the static condition cutoff of _try_factor_mixed_kkt! (mixed_precision_kkt.jl:1049-1059): the
        # Schur matrix's condition estimate from its factor (cond_1(S) as xPOCON estimates it; the
        # reference's xTRCON product of :371-392 bounds it from above and is reported beside it) under
        # the reduced-precision limit.  The predicted-refinement-step budget (:1079-1093) is left to
        # the measured refinement guards below, as the reference's :on mode does (:990-1000).
        cond_est = b.factor_condition(1)[0]
        limit = o.mixed_cond_limit if o.precision == "mixed" else o.bf16x3_cond_limit
        if not cond_est <= limit:
            return dict(status="retry_fp64", reason=f"Schur condition estimate {cond_est:.2e} above the reduced-precision limit",
                        cond_estimate=cond_est)
    # predictor (step.jl:699-726)
    b.predictor_rhs()
    b.solve_direction()
    if use32:
        # predictor guard (mixed_precision_kkt.jl:1475-1586): residual <= 1e-8 max(||r||, 1)
        _, rn = b.exact_residual()
        _, pres = b.refine_exact(1e-8 * max(rn, 1.0), o.mixed_max_refine)
        if not pres <= 1e-8 * max(rn, 1.0):
            return dict(status="retry_fp64", reason=f"predictor refinement stalled at {pres:.2e}")
    b.direction_blocks()
    clock.lap("predictor")
    adaptive = ctrl.strategy == "adaptive" and not ctrl.fallback
    kdim = max(sum(k), 1)
    if adaptive:
        # _affine_predictor_diagnostics! (step.jl:546-586)
        ap, ad = b.fraction_to_boundary(1.0, 1.0)
        comps, affs = b.complementarity(ap, ad)
        uniform = True
        ref_v, ref_d = None, 0
        for l in range(b.L):
            v, dim = float(comps[l]), k[l]
            if dim <= 0:
                continue
            if not math.isfinite(v):
                uniform = False
            elif ref_v is None:
                ref_v, ref_d = v, dim
            elif not (v == ref_v if dim == ref_d else v / dim == ref_v / ref_d):
                uniform = False
    else:
        # _legacy_predictor_diagnostics! (step.jl:588-625)
        ap, ad = 1.0, 1.0
        comps, affs = b.complementarity(1.0, 1.0)
        uniform = False
    comp = float(np.sum(comps))
    mu_g = comp / kdim
    mu_aff = max(0.0, float(np.sum(affs)) / kdim)
    pm, dm, quality = b.factor_quality()
    cur = max(p_res / max(scale_p, 1.0), d_res / max(scale_d, 1.0))
    if ctrl.history:
        row = ctrl.history[-1]
        prev = max(row["primal_residual"] / max(scale_p, 1.0), row["dual_residual"] / max(scale_d, 1.0), EPS)
        achieved = cur / prev
    else:
        achieved = 1.0
    diag = IterationDiagnostics(
        iteration=iteration, primal_residual=p_res / max(scale_p, 1.0), dual_residual=d_res / max(scale_d, 1.0),
        relative_gap=gap_rel, mu=mu_g, mu_aff=mu_aff, affine_primal_step=ap, affine_dual_step=ad,
        previous_primal_step=ctrl.last("primal_step", 1.0), previous_dual_step=ctrl.last("dual_step", 1.0),
        backtracking_count=ctrl.last("backtracking_count", 0),
        regularization=0.0 if reg_attempts <= 0 else SQRT_EPS * 10.0 ** (reg_attempts - 1),
        refinement_count=ctrl.last("refinement_count", 0), factorization_quality=quality,
        achieved_residual_reduction=achieved, primal_psd_margin=pm, dual_psd_margin=dm)
    params = ctrl.select_iteration_parameters(diag)
    clock.lap("complementarity_analysis")
    # corrector (step.jl:846-936)
    if adaptive and not params.fallback:
        b.corrector_rhs(True, params.sigma, mu_g, not uniform, mu, 0.0)
    else:
        b.corrector_rhs(False, 0.0, 0.0, False, mu, o_iter.beta)
    b.solve_direction()
    if adaptive and not params.fallback:
        refine_tol, refine_cap = params.refinement_tolerance, params.refinement_max_count
    else:
        refine_tol, refine_cap = o_iter.refine_tol, o_iter.refine_max_steps
    fixed_policy = o_iter.refine_policy == "fixed"
    if use32:
        # _refine_mixed_direction! (mixed_precision_kkt.jl:1634-1713) to the FP64 target
        _, r_norm = b.exact_residual()
        target = o.mixed_refine_target * max(r_norm, 1.0)
        steps, res = b.refine_exact(target, o.mixed_max_refine)
        if not res <= target:
            return dict(status="retry_fp64", reason=f"corrector refinement stalled at {res:.2e}")
        refine_steps = min(steps, 1)  # the controller sees a native-equivalent count
    else:
        if fixed_policy:
            refine_cap = o_iter.refine_steps
            abstol = 0.0
        else:
            _, r_norm = b.direction_residual()
            rel = refine_tol if refine_tol > 0 else REFINE_DEFAULT_TOL_ULPS * EPS
            abstol = rel * max(r_norm, 1.0)
        refine_steps, _ = b.refine(abstol, refine_cap, fixed_policy)
        res, r_norm = b.direction_residual()
    tol = max(SQRT_EPS, o.eps_primal, o.eps_dual) * max(r_norm, 1.0)
    if not math.isfinite(res) or res > tol:
        return dict(status="breakdown",
                    reason=f"final structured KKT direction residual {res} exceeded the accepted tolerance {tol}",
                    p_res=p_res, d_res=d_res, reg_attempts=reg_attempts)
    b.direction_blocks()
    clock.lap("corrector")
    return dict(status="ok", p_res=p_res, d_res=d_res, params=params, comp=comp, mu=mu_g, mu_aff=mu_aff,
                refine_steps=refine_steps, refine_residual=res, reg_attempts=reg_attempts)


def solve(prob: SDPProblem, o: SolverOptions = SolverOptions(), x0=None, X0=None, Y0=None, device: int = 0,
          timing: bool = False, trace_schur: int = 0, backend: MI355XDenseBackend | None = None,
          start: str = "given", equilibrate: bool = False, ruiz_iters="auto") -> SDPResult:
    """Solve a core-form SDP on the GPU (interior_point.jl:1323-2420).

    start="given" takes x0/X0/Y0 (the warm-start path; Omega-scaled identities when X0 is None);
    start="cold" runs the identity-metric KKT cold start on the device (interior_point.jl:699-1287).
    equilibrate=True scales the device-resident problem first (Ruiz, ingest.jl:1432-1521) and maps
    the result back (unequilibrate, ingest.jl:1523-1560); the objectives are then the original
    problem's.  A backend passed in is scaled in place."""
    own = backend is None
    b = backend if backend is not None else MI355XDenseBackend(prob, device)
    try:
        if o.kkt_route == "block_arrow" and b.arrow_info()[0] < 0:
            b.set_arrow()
        if not equilibrate:
            return _solve(b, prob, o, x0, X0, Y0, timing, trace_schur, start)
        eq = b.equilibrate(ruiz_iters)
        if start != "cold" and X0 is not None:  # warm start into the scaled space (interior_point.jl:606-661)
            x0 = np.asarray(x0) * eq.s
            X0 = [Xl * np.outer(E, E) for Xl, E in zip(X0, eq.E)]
            Y0 = [Yl / np.outer(E, E) / eq.objective_scale for Yl, E in zip(Y0, eq.E)]
        r = _solve(b, b.prob, o, x0, X0, Y0, timing, trace_schur, start)
        x, X, _, Y = eq.unequilibrate(r.x, r.X, r.Y)
        r.x, r.X, r.Y = x, X, Y
        r.pobj = float(prob.c @ x)
        r.dobj = float(sum(np.sum(prob.C[l] * Y[l]) for l in range(prob.L)))
        r.gap_rel = abs(r.pobj - r.dobj) / max(1.0, (abs(r.pobj) + abs(r.dobj)) / 2)
        r.equilibration = eq
        return r
    finally:
        if own:
            b.close()


def _solve(b, prob, o, x0, X0, Y0, timing, trace_schur, start="given") -> SDPResult:
    m, L, k = prob.m, prob.L, prob.k
    x = np.zeros(m) if x0 is None else np.asarray(x0, dtype=np.float64)
    if start == "cold":
        rec = b.cold_start(o.beta, o.eps_primal, o.eps_dual)
        if not rec["ok"]:
            X = [np.eye(kl) for kl in k]
            return SDPResult("NumericalBreakdown", x, X, X, 0.0, 0.0, 0, math.inf, math.inf, math.inf,
                             f"cold start failed: {rec['reason']}")
        x, X, Y = b.get_iterate()
    elif X0 is None:
        X = [o.omega_p * np.eye(kl) for kl in k]
        Y = [o.omega_d * np.eye(kl) for kl in k]
    else:
        X, Y = list(X0), list(Y0)
        for M in X + Y:
            try:
                np.linalg.cholesky(M)
            except np.linalg.LinAlgError:
                return SDPResult("NumericalBreakdown", x, X, Y, 0.0, 0.0, 0, math.inf, math.inf, math.inf,
                                 "initial X0/Y0 must be positive definite")
    if start != "cold":
        b.set_iterate(x, X, Y)
    mu = np.array([o.beta * float(np.sum(X[l] * Y[l])) / k[l] for l in range(L)])
    ctrl = AdaptiveIPMController.from_options(o)
    clock = _Clock(b, timing)
    trace = [] if trace_schur else None
    pobj, dobj = b.objectives()
    p_res, d_res, _ = b.compute_residuals(mu, affine=o.parameter_strategy == "adaptive", factor=False)
    scale_p = 1.0 + max((float(np.max(np.abs(Cl))) if Cl.size else 0.0) for Cl in prob.C)
    scale_d = 1.0 + (float(np.max(np.abs(prob.c))) if prob.c.size else 0.0)
    stag = StagnationDetector(o.stall_iterations, o.stall_tolerance)
    best_merit, initial_merit, best = math.inf, math.inf, None
    comp_cur = float(sum(np.sum(X[l] * Y[l]) for l in range(L)))
    iteration, restarts, centering = 0, 0, 0
    status, message = "NotStarted", ""
    mixed_active = o.precision != "fp64"
    fp32_iters = fp64_iters = 0
    gap_rel = math.inf
    history = []
    while True:
        gap_rel = abs(pobj - dobj) / max(1.0, (abs(pobj) + abs(dobj)) / 2)
        term_ok = p_res / scale_p <= o.eps_primal and d_res / scale_d <= o.eps_dual
        gap_ok = gap_rel <= o.eps_gap
        if term_ok and gap_ok:
            p_res, d_res = b.solution_residuals()
            term_ok = p_res / scale_p <= o.eps_primal and d_res / scale_d <= o.eps_dual
        merit = StagnationDetector.merit(o, p_res, d_res, gap_rel, comp_cur, scale_p, scale_d, max(abs(pobj), abs(dobj)))
        floor = StagnationDetector.at_precision_floor(p_res, d_res, gap_rel, scale_p, scale_d)
        stagnated = stag.observe(merit, floor, o.iter_max - iteration)
        if math.isfinite(merit) and not math.isfinite(initial_merit):
            initial_merit = merit
        if math.isfinite(merit) and merit < best_merit:
            best_merit = merit
            if not (term_ok and gap_ok):
                best = (b.get_iterate(), pobj, dobj, p_res, d_res, gap_rel)
        if term_ok and gap_ok:
            status, message = "Optimal", "Optimal"
            break
        if iteration >= o.iter_max:
            status, message = "IterLimit", f"Cannot reach optimality within {o.iter_max} iterations."
            break
        if stagnated:
            status = "InsufficientPrecision" if stag.reason == "precision_floor" else "Stalled"
            message = f"stagnation: {stag.reason}"
            break
        o_iter = ctrl.iteration_options(o)
        tr = trace if (trace is not None and len(trace) < trace_schur) else None
        res = newton_step(b, o_iter, o, mu, ctrl, iteration + 1, gap_rel, scale_p, scale_d, clock, tr,
                          use32=mixed_active)
        if res["status"] == "retry_fp64":
            # FP32 assembly no longer good enough: redo this step in FP64 and stay there
            mixed_active = False
            res = newton_step(b, o_iter, o, mu, ctrl, iteration + 1, gap_rel, scale_p, scale_d, clock, tr)
        if mixed_active:
            fp32_iters += 1
        else:
            fp64_iters += 1
        if res["status"] == "breakdown":
            status, message = "NumericalBreakdown", res["reason"]
            break
        p_res, d_res = res["p_res"], res["d_res"]
        params = res["params"]
        clock.start()
        rule = resolved_step_rule(k, o.step_rule)
        if rule == "fraction_to_boundary":
            tX, tY = b.fraction_to_boundary(params.primal_fraction_to_boundary, params.dual_fraction_to_boundary)
        else:
            ratio = SQRT_EPS if o.parameter_strategy == "adaptive" else 0.0
            tX, tY = b.backtrack(params.backtracking_factor, params.minimum_step, ratio)
        clock.lap("line_search")
        bcount = (estimate_backtracking_count(tX, params.backtracking_factor, rule)
                  + estimate_backtracking_count(tY, params.backtracking_factor, rule))
        primal_feasible = p_res / scale_p <= o.eps_primal
        dual_feasible = d_res / scale_d <= o.eps_dual
        x_stuck, y_stuck = tX < params.minimum_step, tY < params.minimum_step
        if x_stuck and primal_feasible and not y_stuck:
            tX, x_stuck = 0.0, False
        if y_stuck and dual_feasible and not x_stuck:
            tY, y_stuck = 0.0, False
        if x_stuck or y_stuck:
            # collapse handling (interior_point.jl:2010-2217)
            near = 1000.0
            if gap_rel < near * o.eps_gap and p_res / scale_p < near * o.eps_primal and d_res / scale_d < near * o.eps_dual:
                status, message = "Stalled", "Step size collapsed near the requested tolerance."
                break
            if (o.stall_iterations > 0 and math.isfinite(initial_merit) and best_merit < initial_merit * 1e-3
                    and best_merit < NEAR_SOLUTION_MERIT):
                status, message = "InsufficientPrecision", "Step size collapsed after convergence progress."
                break
            if centering < o.max_centering and ctrl.beta < CENTERING_BETA_MAX:
                centering += 1
                ctrl.beta = min(ctrl.beta * CENTERING_BETA_STEP, CENTERING_BETA_MAX)
                comps, _ = b.complementarity(0.0, 0.0)
                mu = np.array([ctrl.beta * float(comps[l]) / k[l] for l in range(L)])
                continue
            rescale_X = x_stuck and not primal_feasible
            rescale_Y = y_stuck and not dual_feasible
            if o.restart and restarts < o.max_restarts and (rescale_X or rescale_Y):
                restarts += 1
                if rescale_X:
                    b.scale_blocks(1, params.restart_scale)
                if rescale_Y:
                    b.scale_blocks(2, params.restart_scale)
                comps, _ = b.complementarity(0.0, 0.0)
                comp_cur = float(np.sum(comps))
                mu = np.array([ctrl.beta * float(comps[l]) / k[l] for l in range(L)])
                continue
            if o.restart and restarts < o.max_restarts:
                status, message = "Stalled", "Step size collapsed while both residuals were already within tolerance."
            else:
                status = "MaxRestartsExceeded" if restarts >= o.max_restarts else "NumericalBreakdown"
                message = "Step size collapsed."
            break
        comp_after, finite, block_comp = b.update(tX, tY)
        if not (finite and np.all(np.isfinite(mu))):
            status, message = "NumericalBreakdown", "non-finite primal or dual iterate detected"
            break
        # block primal residual = max ||P_l||_inf of this iteration (step.jl:931): p_res here
        p_after, d_after = b.accepted_residuals(tX, tY, res["p_res"])
        ctrl.record_and_update(iteration=iteration + 1, complementarity_before=res["comp"],
                               complementarity_after=comp_after, primal_residual=p_res, dual_residual=d_res,
                               primal_step=tX, dual_step=tY, primal_residual_after=p_after,
                               dual_residual_after=d_after, backtracking_count=bcount,
                               refinement_count=res["refine_steps"], selected=params)
        comp_cur = comp_after
        if ctrl.strategy == "adaptive" and not ctrl.fallback:
            mu = np.full(L, ctrl.beta * comp_after / max(sum(k), 1))
        else:
            mu = np.array([ctrl.beta * float(block_comp[l]) / k[l] for l in range(L)])
        pobj, dobj = b.objectives()
        clock.lap("update")
        iteration += 1
        history.append(dict(iteration=iteration, pobj=pobj, dobj=dobj, p_res=p_after, d_res=d_after, tX=tX, tY=tY,
                            reg_attempts=res["reg_attempts"], refine_steps=res["refine_steps"]))
        if o.verbosity >= 1:
            print(f"{iteration}\t{pobj:.5E}\t{dobj:.5E}\t{pobj - dobj:.5E}\t{p_after:.5E}\t{d_after:.5E}\t"
                  f"{tX:.5E}\t{tY:.5E}", flush=True)
        p_res, d_res = p_after, d_after
    if status != "Optimal" and best is not None:
        (x, X, Y), pobj, dobj, p_res, d_res, gap_rel = best
    else:
        x, X, Y = b.get_iterate()
    return SDPResult(status, x, X, Y, pobj, dobj, iteration, p_res, d_res, gap_rel, message, clock.t, history,
                     trace or [], fp32_iters, fp64_iters)
