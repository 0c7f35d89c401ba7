"""The adaptive interior-point controller against the reference's own known values.

Each case restates a testset of test/adaptive_parameter_policy.jl (inputs and
the values it asserts) and runs it on the three restatements of the
controller in this repository:

* the oracle (oracle/sdpx_oracle.py: AdaptivePolicy, Controller),
* the single-problem route's host controller (sdpx_mi355x/policy.py),
* the batched route's C++ controller (csrc/batch_api.hip), through the host-only
  hooks sdpx_batch_policy_select / sdpx_batch_policy_record (no device call).
"""

import ctypes as C
import math

import numpy as np
import pytest

import sdpx_oracle as O

from sdpx_mi355x import policy as P
from sdpx_mi355x._lib import check, dptr, load
from sdpx_mi355x.batch import BatchOptions

DIAG_FIELDS = ("iteration", "primal_residual", "dual_residual", "relative_gap", "mu", "mu_aff",
               "affine_primal_step", "affine_dual_step", "previous_primal_step", "previous_dual_step",
               "backtracking_count", "regularization", "refinement_count", "factorization_quality",
               "achieved_residual_reduction", "primal_psd_margin", "dual_psd_margin")
REASONS = {0: "none", 1: "nonfinite_diagnostics", 2: "degraded_factorization", 3: "repeated_instability",
           4: "previous_instability"}


def _diag_values(d: dict) -> list:
    base = dict(iteration=0, primal_residual=math.inf, dual_residual=math.inf, relative_gap=math.inf, mu=math.inf,
                mu_aff=math.inf, affine_primal_step=0.0, affine_dual_step=0.0, previous_primal_step=1.0,
                previous_dual_step=1.0, backtracking_count=0, regularization=0.0, refinement_count=0,
                factorization_quality=1.0, achieved_residual_reduction=1.0, primal_psd_margin=1.0,
                dual_psd_margin=1.0)
    base.update(d)
    return base


class Oracle:
    name = "oracle"

    def select(self, opts, diag, history=(), preset_fallback=False):
        c = O.Controller(O.Options(**opts))
        c.history = [dict(sigma=s, unstable=u) for s, u in history]
        c.fallback = preset_fallback
        p = c.select(O.Diagnostics(**_diag_values(diag)))
        o2 = c.iteration_options(O.Options(**opts))
        return dict(sigma=p.sigma, pf=p.primal_fraction_to_boundary, df=p.dual_fraction_to_boundary,
                    bt=p.backtracking_factor, rtol=p.refinement_tolerance, rmax=p.refinement_max_count,
                    fallback=p.fallback, reason=p.fallback_reason, ctrl_fallback=c.fallback, beta=o2.beta,
                    gamma=o2.gamma, strategy=o2.parameter_strategy,
                    sigma_max=getattr(c.policy, "sigma_max", None))

    def record(self, opts, rec):
        c = O.Controller(O.Options(**opts))
        c.record(selected=c.parameters, refinement_count=0, **rec)
        return c.history[-1]["achieved_residual_reduction"], c.history[-1]["unstable"]


class Policy:
    name = "policy.py"

    def select(self, opts, diag, history=(), preset_fallback=False):
        o = P.SolverOptions(**opts)
        c = P.AdaptiveIPMController.from_options(o)
        c.history = [dict(sigma=s, unstable=u) for s, u in history]
        c.fallback = preset_fallback
        p = c.select_iteration_parameters(P.IterationDiagnostics(**_diag_values(diag)))
        o2 = c.iteration_options(o)
        return dict(sigma=p.sigma, pf=p.primal_fraction_to_boundary, df=p.dual_fraction_to_boundary,
                    bt=p.backtracking_factor, rtol=p.refinement_tolerance, rmax=p.refinement_max_count,
                    fallback=p.fallback, reason=p.fallback_reason, ctrl_fallback=c.fallback, beta=o2.beta,
                    gamma=o2.gamma, strategy=o2.parameter_strategy,
                    sigma_max=getattr(c.policy, "sigma_max", None))

    def record(self, opts, rec):
        c = P.AdaptiveIPMController.from_options(P.SolverOptions(**opts))
        c.record_and_update(selected=c.parameters, refinement_count=0, **rec)
        return c.history[-1]["achieved_residual_reduction"], c.history[-1]["unstable"]


class Cpp:
    name = "batch_api.hip"

    def select(self, opts, diag, history=(), preset_fallback=False):
        lib = load()
        bo = BatchOptions.from_solver_options(P.SolverOptions(**opts))
        d = _diag_values(diag)
        dv = np.array([float(d[f]) for f in DIAG_FIELDS])
        hist = np.array([[s, 1.0 if u else 0.0] for s, u in history], dtype=np.float64).ravel()
        out = np.zeros(16)
        check(lib.sdpx_batch_policy_select(C.addressof(bo), dptr(dv), len(history), dptr(hist) if len(history) else None,
                                           int(preset_fallback), dptr(out)), "sdpx_batch_policy_select")
        fb = bool(out[11])
        return dict(sigma=out[0], pf=out[1], df=out[2], bt=out[3], rtol=out[4], rmax=int(out[5]),
                    fallback=bool(out[9]), reason=REASONS[int(out[10])], ctrl_fallback=fb, beta=out[12],
                    gamma=out[13], strategy="fixed" if (fb or not bo.adaptive) else "adaptive",
                    sigma_max=out[14] if bo.adaptive else None)

    def record(self, opts, rec):
        lib = load()
        bo = BatchOptions.from_solver_options(P.SolverOptions(**opts))
        keys = ("iteration", "complementarity_before", "complementarity_after", "primal_residual", "dual_residual",
                "primal_step", "dual_step", "primal_residual_after", "dual_residual_after", "backtracking_count")
        rv = np.array([float(rec[k]) for k in keys] + [0.0])
        out = np.zeros(4)
        check(lib.sdpx_batch_policy_record(C.addressof(bo), dptr(rv), dptr(out)), "sdpx_batch_policy_record")
        return out[0], bool(out[1])


IMPLS = [Oracle(), Policy(), Cpp()]


@pytest.fixture(params=IMPLS, ids=lambda i: i.name)
def impl(request):
    return request.param


def test_fixed_policy_reproduces_configured_values(impl):
    """adaptive_parameter_policy.jl:20-38"""
    opts = dict(beta=0.075, gamma=0.8, refine_tol=1e-13, refine_max_steps=5, parameter_strategy="fixed")
    s = impl.select(opts, {})
    assert s["sigma"] == 0.075
    assert s["pf"] == s["df"] == s["bt"] == 0.8
    assert s["rtol"] == 1e-13 and s["rmax"] == 5
    assert not s["fallback"]


def test_adaptive_selection_is_bounded(impl):
    """adaptive_parameter_policy.jl:40-71 (the Float64 instance)."""
    diag = dict(iteration=2, primal_residual=1 / 1000, dual_residual=1 / 2000, relative_gap=1 / 100, mu=1.0,
                mu_aff=1 / 2, affine_primal_step=9 / 10, affine_dual_step=4 / 5)
    s = impl.select(dict(parameter_strategy="adaptive"), diag)
    assert 1 / 50 <= s["sigma"] <= 1 / 2
    assert 4 / 5 <= s["pf"] <= 99 / 100 and 4 / 5 <= s["df"] <= 99 / 100
    assert not s["fallback"]


def test_short_affine_steps_request_more_centering(impl):
    """adaptive_parameter_policy.jl:73-104"""
    common = dict(iteration=3, primal_residual=1e-3, dual_residual=1e-3, relative_gap=1e-2, mu=1.0, mu_aff=1e-3)
    opts = dict(parameter_strategy="adaptive")
    long_step = impl.select(opts, dict(common, affine_primal_step=0.95, affine_dual_step=0.95))
    short_step = impl.select(opts, dict(common, affine_primal_step=0.01, affine_dual_step=0.02))
    assert short_step["sigma"] > long_step["sigma"]
    assert short_step["pf"] < long_step["pf"]


def test_generic_sigma_cap_prevents_over_centering(impl):
    """adaptive_parameter_policy.jl:106-135: sigma == 0.15 under adaptive_sigma_max = 0.15."""
    opts = dict(beta=0.075, gamma=0.8, adaptive_sigma_max=0.15, parameter_strategy="adaptive")
    diag = dict(iteration=6, primal_residual=0.70, dual_residual=0.006, relative_gap=0.52, mu=0.00152, mu_aff=0.00142,
                affine_primal_step=0.019, affine_dual_step=0.135, previous_primal_step=0.512,
                previous_dual_step=0.640, backtracking_count=5)
    s = impl.select(opts, diag, history=[(0.272, False)])
    assert s["sigma"] == 0.15
    assert s["sigma_max"] == 0.15


@pytest.mark.parametrize("beta,requested,expected", [(0.075, 0.0, 0.5), (0.1, 0.0, 0.5), (0.075, 0.2, 0.2),
                                                     (0.2, 0.1, 0.2)])
def test_recommended_sigma_cap(impl, beta, requested, expected):
    """adaptive_parameter_policy.jl:136-151: recommended_adaptive_sigma_max (interior_point.jl:166-172) and
    AdaptiveParameterPolicy(beta=0.2, adaptive_sigma_max=0.1).sigma_max == 0.2."""
    s = impl.select(dict(beta=beta, adaptive_sigma_max=requested, parameter_strategy="adaptive"), {})
    assert s["sigma_max"] == expected


def test_nonfinite_diagnostics_fall_back_explicitly(impl):
    """adaptive_parameter_policy.jl:169-195"""
    opts = dict(beta=0.075, gamma=0.8, parameter_strategy="adaptive")
    diag = dict(iteration=4, primal_residual=math.nan, dual_residual=1.0, relative_gap=1.0, mu=1.0, mu_aff=0.5,
                affine_primal_step=0.5, affine_dual_step=0.5)
    s = impl.select(opts, diag)
    assert s["fallback"] and s["reason"] == "nonfinite_diagnostics"
    assert s["sigma"] == 0.075 and s["bt"] == 0.8


def test_degraded_factorization_restores_fixed_policy(impl):
    """adaptive_parameter_policy.jl:197-230"""
    opts = dict(beta=0.075, gamma=0.8, parameter_strategy="adaptive")
    diag = dict(iteration=5, primal_residual=1e-5, dual_residual=1e-5, relative_gap=1e-4, mu=1e-3, mu_aff=1e-4,
                affine_primal_step=0.8, affine_dual_step=0.8, factorization_quality=0.0)
    s = impl.select(opts, diag)
    assert s["fallback"] and s["reason"] == "degraded_factorization"
    assert s["ctrl_fallback"]
    assert s["strategy"] == "fixed" and s["beta"] == 0.075 and s["gamma"] == 0.8


def test_sigma_three_twenty_fifths(impl):
    """adaptive_parameter_policy.jl:232-256 pins sigma == 3/25 in BigFloat (256 bits).  In Float64 the
    same selection is one rounding of each of its operations away: within 2 ulp of 3/25."""
    diag = dict(iteration=2, primal_residual=1 / 1000, dual_residual=1 / 1000, relative_gap=1 / 100, mu=1.0,
                mu_aff=1 / 2, affine_primal_step=9 / 10, affine_dual_step=4 / 5)
    s = impl.select(dict(parameter_strategy="adaptive"), diag)
    assert abs(s["sigma"] - 3 / 25) <= 2 * math.ulp(3 / 25)
    assert not s["fallback"]


def test_previous_instability_keeps_fixed_values(impl):
    """select_iteration_parameters! of a controller already in fallback (adaptive_parameters.jl:511-548)."""
    opts = dict(beta=0.075, gamma=0.8, parameter_strategy="adaptive")
    diag = dict(iteration=7, primal_residual=1e-5, dual_residual=1e-5, relative_gap=1e-4, mu=1e-3, mu_aff=1e-4,
                affine_primal_step=0.8, affine_dual_step=0.8)
    s = impl.select(opts, diag, preset_fallback=True)
    assert s["fallback"] and s["reason"] == "previous_instability"
    assert s["sigma"] == 0.075 and s["bt"] == 0.8 and s["beta"] == 0.075 and s["gamma"] == 0.8


def test_repeated_instability_falls_back(impl):
    """_recent_instability_count >= instability_limit (2) (adaptive_parameters.jl:202-209, 277-283)."""
    opts = dict(parameter_strategy="adaptive")
    diag = dict(iteration=8, primal_residual=1e-5, dual_residual=1e-5, relative_gap=1e-4, mu=1e-3, mu_aff=1e-4,
                affine_primal_step=0.8, affine_dual_step=0.8)
    s = impl.select(opts, diag, history=[(0.2, False), (0.3, True), (0.3, True)])
    assert s["fallback"] and s["reason"] == "repeated_instability"
    s = impl.select(opts, diag, history=[(0.3, True), (0.3, False)])
    assert not s["fallback"]


def test_controller_records_achieved_reduction(impl):
    """adaptive_parameter_policy.jl:260-285: achieved_residual_reduction == 0.75."""
    rec = dict(iteration=1, complementarity_before=1.0, complementarity_after=0.4, primal_residual=1.0,
               dual_residual=0.5, primal_step=1.0, dual_step=1.0, primal_residual_after=0.75,
               dual_residual_after=0.25, backtracking_count=0)
    achieved, unstable = impl.record(dict(parameter_strategy="adaptive"), rec)
    assert achieved == 0.75 and not unstable


def test_feasibility_growth_is_unstable(impl):
    """_outcome_is_unstable (adaptive_parameters.jl:550-586): feasibility ratio above 3 after iteration 3
    is unstable; on a fresh controller the ratio is 1 (no previous feasibility)."""
    rec = dict(iteration=5, complementarity_before=1.0, complementarity_after=0.5, primal_residual=4.0,
               dual_residual=1.0, primal_step=1.0, dual_step=1.0, primal_residual_after=1.0,
               dual_residual_after=1.0, backtracking_count=0)
    _, unstable = impl.record(dict(parameter_strategy="adaptive"), rec)
    assert not unstable
    rec.update(complementarity_after=2.5)  # complementarity growth above 2
    _, unstable = impl.record(dict(parameter_strategy="adaptive"), rec)
    assert unstable
