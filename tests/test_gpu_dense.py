"""GPU parity of the dense Newton-step phases against the CPU oracle (through the C-ABI)."""

import math

import numpy as np
import pytest

import sdpx_oracle as O
from sdpx_mi355x import MI355XDenseBackend, SolverOptions, ingest, solve
from sdpx_mi355x import generators as G

import _helpers as _H

pytestmark = pytest.mark.gpu


def _rel_fro_lower(A, B):
    La, Lb = np.tril(A), np.tril(B)
    return np.linalg.norm(La - Lb) / max(np.linalg.norm(Lb), 1e-300)


def _random_pd(rng, k, cond=10.0):
    Q, _ = np.linalg.qr(rng.standard_normal((k, k)))
    ev = np.exp(rng.uniform(0, math.log(cond), k))
    return (Q * ev) @ Q.T


def _dense_case(k, m, seed):
    g = G.dense_feasible(k, m, seed=seed)
    prob = ingest(g.c, [g.A_blocks()], [g.C])
    oprob = O.ingest(g.c, [g.A_blocks()], [g.C])
    return g, prob, oprob


@pytest.mark.parametrize("k,m", [(7, 11), (50, 200), (65, 130), (130, 300)])
def test_schur_build_matches_oracle(gpu_lib, k, m):
    rng = np.random.default_rng(k * 1000 + m)
    g, prob, oprob = _dense_case(k, m, seed=3)
    X, Y = _random_pd(rng, k), _random_pd(rng, k)
    b = MI355XDenseBackend(prob)
    b.set_iterate(g.x0, [X], [Y])
    _, _, ok = b.compute_residuals([0.1], affine=True, factor=True)
    assert ok
    b.schur_build()
    S = b.get_schur()
    LX, _ = O.kchol(X)
    MY, _ = O.kchol(Y)
    Sref = O.schur_build(oprob, [LX], [MY])
    assert _rel_fro_lower(S, Sref) <= 1e-12
    b.close()


@pytest.mark.parametrize("k,m", [(65, 130), (130, 700)])
def test_autotuned_schur_matches_oracle(gpu_lib, k, m):
    """The autotuner's tile and split choices leave the Schur matrix within the FP64 bar."""
    rng = np.random.default_rng(k + m)
    g, prob, oprob = _dense_case(k, m, seed=4)
    X, Y = _random_pd(rng, k), _random_pd(rng, k)
    b = MI355XDenseBackend(prob)
    default_split = b.schur_split()
    t = b.autotune()  # before the iterate
    print(t, flush=True)
    assert all(sh in ("128x128", "64x64", "128x64", "64x128") for sh in t["panel_tiles"])
    assert t["tile_group"] in (4, 8, 16)
    for p, cand in enumerate(t["panel_candidates_ms"]):
        # a candidate replaces the incumbent only when 1% faster
        assert cand[t["panel_tiles"][p]] <= min(cand.values()) / 0.99 + 1e-3
    assert t["default_split"] == default_split
    assert t["split"] in {default_split, max(1, 3 * default_split // 4), max(1, default_split // 2),
                          3 * default_split // 2, 2 * default_split}
    assert b.schur_split() == t["split"] and t["panel_ms"] > 0 and t["gram_ms"] > 0
    b.set_iterate(g.x0, [X], [Y])
    _, _, ok = b.compute_residuals([0.1], affine=True, factor=True)
    assert ok
    b.schur_build()
    S = b.get_schur()
    LX, _ = O.kchol(X)
    MY, _ = O.kchol(Y)
    assert _rel_fro_lower(S, O.schur_build(oprob, [LX], [MY])) <= 1e-12
    b.close()


def test_autotune_restores_the_iterate_factors(gpu_lib):
    """The autotuner times its candidates with stand-ins for L_X^-1 and M_Y and puts
    the iterate's back: with the earlier tiling re-applied, a Schur build from the
    same iterate (no new residual pass) is bitwise the one before tuning."""
    k, m = 130, 300
    rng = np.random.default_rng(12)
    g, prob, _ = _dense_case(k, m, seed=7)
    X, Y = _random_pd(rng, k), _random_pd(rng, k)
    b = MI355XDenseBackend(prob)
    try:
        b.set_tiling(("64x64", "64x64"), split=b.schur_split(), tile_group=8)
        split = b.schur_split()
        b.set_iterate(g.x0, [X], [Y])
        _, _, ok = b.compute_residuals([0.1], affine=True, factor=True)
        assert ok
        b.schur_build()
        S0 = np.tril(b.get_schur())
        b.autotune()
        b.set_tiling(("64x64", "64x64"), split=split, tile_group=8)
        b.schur_build()
        assert np.array_equal(np.tril(b.get_schur()), S0)
    finally:
        b.close()


@pytest.mark.parametrize("shapes", [("64x64", "64x64"), ("128x64", "64x128"), ("64x128", "128x64"),
                                    ("128x128", "128x64")])
def test_panel_tile_shapes_match_oracle(gpu_lib, shapes):
    """Every panel tile shape the autotuner may pick gives the Schur matrix within the FP64 bar
    (k = 130: partial tiles on both sides and triangular K ranges)."""
    k, m = 130, 300
    rng = np.random.default_rng(11)
    g, prob, oprob = _dense_case(k, m, seed=6)
    X, Y = _random_pd(rng, k), _random_pd(rng, k)
    b = MI355XDenseBackend(prob)
    b.set_tiling(shapes, split=1, tile_group=4)
    assert b.schur_split() == 1
    b.set_iterate(g.x0, [X], [Y])
    _, _, ok = b.compute_residuals([0.1], affine=True, factor=True)
    assert ok
    b.schur_build()
    S = b.get_schur()
    LX, _ = O.kchol(X)
    MY, _ = O.kchol(Y)
    assert _rel_fro_lower(S, O.schur_build(oprob, [LX], [MY])) <= 1e-12
    with pytest.raises(Exception):
        b.set_tiling(shapes, split=10 ** 6)
    b.close()


def test_schur_identity_is_gram_and_upper_untouched(gpu_lib):
    k, m = 40, 150
    g, prob, _ = _dense_case(k, m, seed=5)
    b = MI355XDenseBackend(prob)
    sentinel = np.full((m, m), -17.0)
    b.set_schur(sentinel)
    b.set_iterate(g.x0, [np.eye(k)], [np.eye(k)])
    b.compute_residuals([0.1], affine=True, factor=True)
    b.schur_build()
    S = b.get_schur()
    A = g.Av_rows
    gram = A @ A.T
    assert _rel_fro_lower(S, gram) <= 1e-13
    iu = np.triu_indices(m, 1)
    assert np.all(S[iu] == -17.0), "upper triangle must not be written"
    # determinism: a second build is bitwise identical
    b.schur_build()
    S2 = b.get_schur()
    assert np.array_equal(np.tril(S), np.tril(S2))
    b.close()


def test_factor_solve_refine(gpu_lib):
    k, m = 30, 257
    rng = np.random.default_rng(0)
    g, prob, _ = _dense_case(k, m, seed=1)
    b = MI355XDenseBackend(prob)
    M = rng.standard_normal((m, m))
    S = M @ M.T + m * np.eye(m)
    b.set_schur(S)
    ok, attempts = b.factorize()
    assert ok and attempts == 0
    r = rng.standard_normal(m)
    b.set_rhs(r)
    b.solve_direction()
    dx, _, _ = b.get_direction()
    assert np.linalg.norm(S @ dx - r, np.inf) <= 1e-12 * np.linalg.norm(r, np.inf) * m
    res, rn = b.direction_residual()
    assert res <= 1e-11 and abs(rn - np.max(np.abs(r))) == 0.0
    steps, res2 = b.refine(64 * 2.2e-16 * max(rn, 1.0), 8, False)
    assert res2 <= res
    b.close()


def test_regularization_contract(gpu_lib):
    """kkt_regressions.jl:927-1000: PD never shifted, -1e-6 repaired, -1 fails after 6."""
    m = 40
    rng = np.random.default_rng(2)
    g, prob, _ = _dense_case(4, m, seed=2)
    b = MI355XDenseBackend(prob)
    Q, _ = np.linalg.qr(rng.standard_normal((m, m)))
    base = np.linspace(1.0, 2.0, m)
    b.set_schur((Q * base) @ Q.T)
    assert b.factorize() == (True, 0)
    ev = base.copy()
    ev[0] = -1e-6
    b.set_schur((Q * ev) @ Q.T)
    ok, attempts = b.factorize()
    assert ok and 1 <= attempts <= 6
    Lo, ok_o, att_o = O.factor_dense_kkt((Q * ev) @ Q.T)
    assert ok_o and att_o == attempts
    ev[0] = -1.0
    b.set_schur((Q * ev) @ Q.T)
    assert b.factorize() == (False, 6)
    b.close()


def test_step_phases_match_oracle(gpu_lib):
    """One full predictor/corrector step: residuals, rhs, directions, line search vs the oracle."""
    k, m = 24, 60
    g, prob, oprob = _dense_case(k, m, seed=11)
    rng = np.random.default_rng(5)
    X, Y = _random_pd(rng, k, 5.0), _random_pd(rng, k, 5.0)
    x = g.x0 + 0.01 * rng.standard_normal(m)
    mu = [0.1 * float(np.sum(X * Y)) / k]
    b = MI355XDenseBackend(prob)
    b.set_iterate(x, [X], [Y])
    p_res, d_res, ok = b.compute_residuals(mu, affine=False, factor=True)
    st = O.StepState()
    o = O.Options(parameter_strategy="fixed")
    p_o, d_o, ok_o = O.compute_residuals(oprob, st, x, [X], [Y], mu, o, factor=True)
    assert ok and ok_o
    assert abs(p_res - p_o) <= 1e-13 * max(1, p_o) and abs(d_res - d_o) <= 1e-13 * max(1, d_o)
    b.schur_build()
    assert b.factorize() == (True, 0)
    st.S = O.schur_build(oprob, st.LX, st.MY)
    st.Lf, _, _ = O.factor_dense_kkt(st.S)
    b.predictor_rhs()
    r = b.get_rhs()
    r_o = -(st.d + O.rhs_v(oprob, st, [Y]))
    assert np.max(np.abs(r - r_o)) <= 1e-11 * max(1.0, np.max(np.abs(r_o)))
    b.solve_direction()
    st.dx = O.solve_dense_kkt(st.Lf, r_o)
    b.direction_blocks()
    dx, dX, dY = b.get_direction()
    O.direction_blocks(oprob, st, [Y])
    assert np.max(np.abs(dx - st.dx)) <= 1e-9 * max(1.0, np.max(np.abs(st.dx)))
    assert np.max(np.abs(dX[0] - st.dX[0])) <= 1e-9 * max(1.0, np.max(np.abs(st.dX[0])))
    assert np.max(np.abs(dY[0] - st.dY[0])) <= 1e-9 * max(1.0, np.max(np.abs(st.dY[0])))
    assert np.array_equal(dY[0], dY[0].T)
    # fixed corrector
    b.corrector_rhs(False, 0.0, 0.0, False, mu, 0.1)
    O.corrector_R_fixed(oprob, st, [X], [Y], mu, 0.1)
    r2 = b.get_rhs()
    r2_o = -(st.d + O.rhs_v(oprob, st, [Y]))
    assert np.max(np.abs(r2 - r2_o)) <= 1e-10 * max(1.0, np.max(np.abs(r2_o)))
    # line searches on the recovered direction
    b.solve_direction()
    b.direction_blocks()
    st.dx = O.solve_dense_kkt(st.Lf, r2_o)
    O.direction_blocks(oprob, st, [Y])
    tX, tY = b.backtrack(0.9, 1e-10, 0.0)
    tX_o, tY_o = O.backtracking_line_search(oprob, [X], [Y], st.dX, st.dY, 0.9, 1e-10, 0.0)
    assert (tX, tY) == (tX_o, tY_o)
    fX, fY = b.fraction_to_boundary(1.0, 1.0)
    fX_o, fY_o = O.fraction_line_search(oprob, [X], [Y], st.dX, st.dY, 1.0, 1.0)
    assert abs(fX - fX_o) <= 1e-12 and abs(fY - fY_o) <= 1e-12
    # the levels per batched trial round only regroup the same sequential bisection
    for levels in (1, 3, 4, 6):
        b.set_tiling(("128x128", "128x128"), split=1, tile_group=8, ftb_levels_per_round=levels)
        assert b.fraction_to_boundary(1.0, 1.0) == (fX, fY), levels
    b.close()


def test_t1_known_answer(gpu_lib):
    """test/correctness.jl:71-94, 230-243: p* = 2 sqrt 6, 19 +- 2 iterations (fixed policy)."""
    A = np.zeros((2, 2, 2))
    A[0, 0, 0] = 1
    A[1, 1, 1] = 1
    Cm = np.array([[0.0, 1.0], [1.0, 0.0]])
    prob = ingest([2.0, 3.0], [A], [Cm])
    res = solve(prob, SolverOptions(parameter_strategy="fixed"))
    assert res.status == "Optimal"
    assert abs(res.pobj - 2 * math.sqrt(6)) < 1e-8
    assert abs(res.iterations - 19) <= 2


@pytest.mark.parametrize("k,m,strategy", [(50, 200, "adaptive"), (50, 200, "fixed"), (32, 128, "adaptive")])
def test_end_to_end_matches_oracle(gpu_lib, k, m, strategy):
    """Plain FP64 GPU path vs the oracle: status, objective 1e-8, certificate, iterations +-2."""
    g, prob, oprob = _dense_case(k, m, seed=0)
    tol = 1e-8
    o = SolverOptions(eps_gap=tol, eps_primal=tol, eps_dual=tol, parameter_strategy=strategy)
    res = solve(prob, o, x0=g.x0, X0=[g.X0], Y0=[g.Y0])
    oo = O.Options(eps_gap=tol, eps_primal=tol, eps_dual=tol, parameter_strategy=strategy)
    ref = O.solve(oprob, oo, x0=g.x0, X0=[g.X0], Y0=[g.Y0])
    assert ref.status == "Optimal"
    assert res.status == ref.status
    assert abs(res.pobj - ref.pobj) <= 1e-8 * max(1.0, abs(ref.pobj))
    assert abs(res.iterations - ref.iterations) <= 2
    cert = O.certificate(oprob, O.Result(res.status, res.x, res.X, res.Y, res.pobj, res.dobj, res.iterations,
                                         res.p_res, res.d_res, res.gap_rel), tol, tol, tol)
    assert cert["valid"], cert


def test_multi_block(gpu_lib):
    """L = 3 blocks of different sizes (the shape of correctness.jl T4)."""
    rng = np.random.default_rng(4)
    m, ks = 30, [5, 12, 3]
    A, Cs = [], []
    y0 = rng.standard_normal(m)
    for k in ks:
        Ab = rng.standard_normal((m, k, k))
        Ab = (Ab + Ab.transpose(0, 2, 1)) / 2
        A.append(Ab)
        Cs.append(np.tensordot(y0, Ab, axes=1) - np.eye(k))
    c = np.array([sum(np.trace(A[l][i]) for l in range(3)) for i in range(m)])
    prob = ingest(c, A, Cs)
    oprob = O.ingest(c, A, Cs)
    o = SolverOptions(eps_gap=1e-8, eps_primal=1e-8, eps_dual=1e-8)
    X0 = [np.eye(k) for k in ks]
    res = solve(prob, o, x0=y0, X0=X0, Y0=X0)
    ref = O.solve(oprob, O.Options(eps_gap=1e-8, eps_primal=1e-8, eps_dual=1e-8), x0=y0, X0=X0, Y0=X0)
    assert res.status == ref.status == "Optimal"
    assert abs(res.pobj - ref.pobj) <= 1e-8 * max(1.0, abs(ref.pobj))
    assert abs(res.iterations - ref.iterations) <= 2


@pytest.mark.parametrize("name,rho,mode,fixed,cap,expected", _H.REFINE_CASES)
def test_refine_matches_oracle(gpu_lib, name, rho, mode, fixed, cap, expected):
    """refine! (kkt_backend.jl:476-485) -> _refine_native_direction! (kkt.jl:3891-3963): steps, final
    residual and refined dx against the oracle's refine_direction, on a factor of S applied to
    M = S + rho G (a chosen contraction per correction, so no decision sits at rounding level)."""
    S, M, r = _H.refinement_case(rho, mode)
    m = S.shape[0]
    _, prob, _ = _dense_case(4, m, seed=2)
    b = MI355XDenseBackend(prob)
    try:
        b.set_schur(S)
        assert b.factorize() == (True, 0)
        b.set_schur(M)  # the factor stays that of S
        b.set_rhs(r)
        b.solve_direction()
        rn = float(np.max(np.abs(r)))
        steps, res = b.refine(64 * O.EPS * max(rn, 1.0), cap, fixed)
        dx, _, _ = b.get_direction()
    finally:
        b.close()
    Lf, _, _ = O.factor_dense_kkt(S)
    o = O.Options(refine_policy="fixed" if fixed else "auto", refine_steps=cap, refine_max_steps=cap)
    dx_o, steps_o, res_o = O.refine_direction(M, Lf, r, O.solve_dense_kkt(Lf, r), o)
    print(f"{name}: gpu steps {steps} res {res:.6e} | oracle steps {steps_o} res {res_o:.6e}", flush=True)
    assert steps == steps_o == expected
    assert abs(res - res_o) <= 1e-8 * res_o + 1e-14 * max(rn, 1.0)
    assert np.max(np.abs(dx - dx_o)) <= 1e-12 * max(1.0, np.max(np.abs(dx_o)))


def test_diagnostics_match_oracle(gpu_lib):
    """a10 phase by phase on a two-block problem: per-block complementarity <X_l, Y_l> and the trial
    <X_l + tX dX_l, Y_l + tY dY_l> (_predictor_complementarity_diagnostics!, step.jl:373-544; the
    oracle's complementarity_diagnostics), solution_residuals (validation.jl:15-58) and the objectives
    (interior_point.jl:7-14)."""
    m = 48
    rng = np.random.default_rng(8)
    g1, g2 = G.dense_feasible(9, m, seed=31), G.dense_feasible(14, m, seed=32)
    A, Cs = [g1.A_blocks(), g2.A_blocks()], [g1.C, g2.C]
    prob, oprob = ingest(g1.c, A, Cs), O.ingest(g1.c, A, Cs)
    X = [_random_pd(rng, 9, 5.0), _random_pd(rng, 14, 5.0)]
    Y = [_random_pd(rng, 9, 5.0), _random_pd(rng, 14, 5.0)]
    x = g1.x0 + 0.01 * rng.standard_normal(m)
    mu = [0.1 * float(np.sum(X[l] * Y[l])) / X[l].shape[0] for l in range(2)]
    b = MI355XDenseBackend(prob)
    try:
        b.set_iterate(x, X, Y)
        b.compute_residuals(mu, affine=True, factor=True)
        b.schur_build()
        assert b.factorize() == (True, 0)
        b.predictor_rhs()
        b.solve_direction()
        b.direction_blocks()
        _, dX, dY = b.get_direction()
        st = O.StepState()
        o = O.Options()
        O.compute_residuals(oprob, st, x, X, Y, mu, o, factor=True)
        st.dX, st.dY = dX, dY  # the GPU's own direction: only the diagnostics are compared
        for tX, tY in ((1.0, 1.0), (0.37, 0.81), (0.0, 0.5)):
            comp, trial = b.complementarity(tX, tY)
            comp_o, aff_o, _ = O.complementarity_diagnostics(oprob, st, X, Y, tX, tY, False)
            blocks = [float(np.sum(X[l] * Y[l])) for l in range(2)]
            trial_ref = [float(np.sum((X[l] + tX * dX[l]) * (Y[l] + tY * dY[l]))) for l in range(2)]
            print(f"t=({tX},{tY}): gpu {comp} {trial} | oracle {comp_o} {aff_o}", flush=True)
            assert np.allclose(comp, blocks, rtol=1e-13, atol=0)
            assert abs(float(np.sum(comp)) - comp_o) <= 1e-13 * abs(comp_o)
            assert abs(float(np.sum(trial)) - aff_o) <= 1e-12 * max(abs(aff_o), float(np.sum(comp)))
            assert np.allclose(trial, trial_ref, rtol=1e-12, atol=1e-12 * max(blocks))
        p_res, d_res = b.solution_residuals()
        p_o, d_o = O.solution_residuals(oprob, x, X, Y)
        assert abs(p_res - p_o) <= 1e-13 * max(1.0, p_o) and abs(d_res - d_o) <= 1e-13 * max(1.0, d_o)
        pobj, dobj = b.objectives()
        assert abs(pobj - float(oprob.c @ x)) <= 1e-13 * max(1.0, abs(pobj))
        assert abs(dobj - O.dual_objective(oprob, Y)) <= 1e-13 * max(1.0, abs(dobj))
    finally:
        b.close()


def _scaled_two_block(seed=1):
    """Two blocks (k = 12 and 7) with rows scaled by powers of ten: Ruiz has work to do in each."""
    rng = np.random.default_rng(seed)
    m = 30
    g1, g2 = G.dense_feasible(12, m, seed=seed), G.dense_feasible(7, m, seed=seed + 50)
    d1, d2 = 10.0 ** rng.integers(-3, 4, 12), 10.0 ** rng.integers(-2, 3, 7)
    A = [g1.A_blocks() * np.outer(d1, d1)[None], g2.A_blocks() * np.outer(d2, d2)[None]]
    Cs = [g1.C * np.outer(d1, d1), g2.C * np.outer(d2, d2) - 0.5 * np.eye(7)]
    c = g1.c * 10.0 ** rng.integers(-2, 3, m)
    return ingest(c, A, Cs), O.ingest(c, A, Cs)


@pytest.mark.parametrize("ruiz_iters", ["auto", 3, 0])
def test_equilibrate_matches_oracle_bitwise(gpu_lib, ruiz_iters):
    """equilibrate (ingest.jl:1432-1521) on the device: maxima, 1/sqrt, products of two factors and
    divisions are correctly rounded, so E, s, the objective scale and the scaled C and c equal the
    oracle's bit for bit; the scaled panels are compared through the Gram (the Schur matrix at X = Y = I)."""
    prob, oprob = _scaled_two_block()
    b = MI355XDenseBackend(prob)
    try:
        eq = b.equilibrate(ruiz_iters)
        sp, oeq = O.equilibrate(oprob, ruiz_iters)
        print(f"passes gpu {eq.passes} oracle {oeq.passes}, objective scale {eq.objective_scale}", flush=True)
        assert eq.passes == oeq.passes
        for El, Eo in zip(eq.E, oeq.E):
            assert np.array_equal(El, Eo)
        assert np.array_equal(eq.s, oeq.s) and eq.objective_scale == oeq.objective_scale
        assert np.array_equal(b.prob.c, sp.c)
        for Cl, Co in zip(b.prob.C, sp.C):
            assert np.array_equal(Cl, Co)
        eyes = [np.eye(12), np.eye(7)]
        b.set_iterate(np.zeros(prob.m), eyes, eyes)
        b.compute_residuals([0.0, 0.0], affine=True, factor=True)
        b.schur_build()
        gram = sum(Av.T @ Av for Av in sp.Av)
        assert _rel_fro_lower(b.get_schur(), gram) <= 1e-14
    finally:
        b.close()


def test_cold_start_matches_oracle(gpu_lib):
    """_kkt_cold_start_initialization (interior_point.jl:699-1287) on the device against the oracle:
    the identity-metric solves, the residual gate and the shifts give the same start."""
    prob, oprob = _scaled_two_block(seed=3)
    o = O.Options()
    b = MI355XDenseBackend(prob)
    try:
        b.equilibrate()
        sp, _ = O.equilibrate(oprob)
        rec = b.cold_start(o.beta, o.eps_primal, o.eps_dual)
        x, X, Y = b.get_iterate()
    finally:
        b.close()
    ok, reason, xo, Xo, Yo, muo, reco = O.cold_start(sp, o)
    print(rec, reco, flush=True)
    assert rec["ok"] and ok and rec["reason"] == reason == "none"
    assert rec["regularization_attempts"] == reco["regularization_attempts"]
    assert np.max(np.abs(x - xo)) <= 1e-11 * max(1.0, np.max(np.abs(xo)))
    for l in range(2):
        assert np.max(np.abs(X[l] - Xo[l])) <= 1e-11 * max(1.0, np.max(np.abs(Xo[l])))
        assert np.max(np.abs(Y[l] - Yo[l])) <= 1e-11 * max(1.0, np.max(np.abs(Yo[l])))
    assert np.allclose(rec["mu"], muo, rtol=1e-11)
    assert abs(rec["largest_primal_shift"] - reco["largest_primal_shift"]) <= 1e-11 * reco["largest_primal_shift"]


def test_equilibrated_cold_solve_matches_oracle(gpu_lib):
    """The reference's default SDP setup on the device (Ruiz, then the identity-metric cold start) and
    the solve, mapped back to the original problem, against the oracle doing the same: status,
    objective 1e-8 relative, iterations +-2, and the certificate on the original problem."""
    prob, oprob = _scaled_two_block(seed=5)
    tol = 1e-8
    o = SolverOptions(eps_gap=tol, eps_primal=tol, eps_dual=tol)
    oo = O.Options(eps_gap=tol, eps_primal=tol, eps_dual=tol)
    res = solve(prob, o, start="cold", equilibrate=True)
    ref, _ = O.solve_equilibrated(oprob, oo, start="cold")
    print(f"gpu {res.status} it={res.iterations} pobj={res.pobj:.12e} | oracle {ref.status} it={ref.iterations} "
          f"pobj={ref.pobj:.12e}", flush=True)
    assert res.status == ref.status == "Optimal"
    assert abs(res.iterations - ref.iterations) <= 2
    assert abs(res.pobj - ref.pobj) <= 1e-8 * max(1.0, abs(ref.pobj))
    cert = O.certificate(oprob, O.Result(res.status, res.x, res.X, res.Y, res.pobj, res.dobj, res.iterations,
                                         0.0, 0.0, res.gap_rel), 1e-7, 1e-7, 1e-7)
    assert cert["valid"], cert


@pytest.mark.parametrize("m,n", [(200, 7), (1030, 130)])
def test_equality_kkt_matches_oracle(gpu_lib, m, n):
    """n_eq > 0 (kkt.jl:1466-1560, 2570-2595) on the device: B~ D, Q, the rank gate and the bordered
    solve against the oracle (which is pinned by a direct solve of [S -B; B' 0])."""
    rng = np.random.default_rng(m + n)
    M = rng.standard_normal((m, m))
    S = M @ M.T / m + np.eye(m)
    B = rng.standard_normal((m, n)) * 10.0 ** rng.integers(-3, 4, n)[None, :]
    r, p = rng.standard_normal(m), rng.standard_normal(n)
    _, prob, _ = _dense_case(3, m, seed=2)
    b = MI355XDenseBackend(prob)
    try:
        b.set_schur(S)
        assert b.factorize() == (True, 0)
        b.set_equalities(B)
        st, mn, mx, tau = b.equality_factor(1e-8)
        Bh, Q, D = b.get_equality()
        b.set_rhs(r)
        dy = b.solve_equality(p)
        dx, _, _ = b.get_direction()
        # a duplicated equality column is refused by the rank gate
        B2 = B.copy()
        B2[:, n - 1] = B2[:, 0]
        b2 = MI355XDenseBackend(prob)
        try:
            b2.set_schur(S)
            b2.factorize()
            b2.set_equalities(B2)
            st2 = b2.equality_factor(1e-8)[0]
        finally:
            b2.close()
    finally:
        b.close()
    Lf, _, _ = O.factor_dense_kkt(S)
    sto, Bho, Do, LQo, tauo = O.factor_equality(Lf, B, 1e-8)
    dxo, dyo = O.solve_equality(Lf, Bho, Do, LQo, r, p)
    print(f"m={m} n={n}: status {st} (oracle {sto}), tau {tau:.3e}; |Bh| rel "
          f"{np.linalg.norm(Bh - Bho) / np.linalg.norm(Bho):.2e}; dx rel {np.linalg.norm(dx - dxo) / np.linalg.norm(dxo):.2e}"
          f" dy rel {np.linalg.norm(dy - dyo) / np.linalg.norm(dyo):.2e}; duplicate -> {st2}", flush=True)
    assert st == sto == 0 and tau == tauo
    assert np.array_equal(D, Do) or np.allclose(D, Do, rtol=1e-13)
    assert np.linalg.norm(Bh - Bho) <= 1e-12 * np.linalg.norm(Bho)
    assert np.linalg.norm(np.tril(Q) - np.tril(Bho.T @ Bho)) <= 1e-12 * np.linalg.norm(Bho.T @ Bho)
    assert np.linalg.norm(dx - dxo) <= 1e-10 * np.linalg.norm(dxo)
    assert np.linalg.norm(dy - dyo) <= 1e-10 * np.linalg.norm(dyo)
    assert np.max(np.abs(S @ dx - B @ dy - r)) <= 1e-9 * np.max(np.abs(r))
    assert np.max(np.abs(B.T @ dx - p)) <= 1e-9 * max(1.0, np.max(np.abs(p)))
    # exactly singular Q: refused on both sides.  Whether POTRF fails (1) or rounds the zero pivot up
    # and the rank gate refuses it (2) depends on rounding (kkt.jl:168-171 anticipates exactly this)
    assert st2 in (1, 2) and O.factor_equality(Lf, B2, 1e-8)[0] in (1, 2)
