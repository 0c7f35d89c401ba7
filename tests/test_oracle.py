"""Pins of the CPU oracle against the reference's known answers, closed forms and brute force."""

import math

import numpy as np
import pytest
import scipy.linalg as sla

import sdpx_oracle as O
import _helpers as _H
from _helpers import golden, owner_problem, random_pd, regularization_matrix, rel_fro_lower
from sdpx_mi355x import generators as G


def test_t1_known_answer_and_iteration_pin():
    """test/correctness.jl:71-94 and the T6 pin (19 iterations, fixed policy)."""
    g = golden()
    prob = O.ingest(g["t1_c"], [g["t1_A"]], [g["t1_C"]])
    r = O.solve(prob, O.Options(parameter_strategy="fixed"))
    assert r.status == "Optimal"
    assert abs(r.pobj - float(g["t1_popt"])) < 1e-8
    assert abs(r.pobj - r.dobj) < 1e-8
    assert abs(r.iterations - int(g["t1_iterations_fixed"])) <= 2
    assert all(np.all(np.linalg.eigvalsh(b + 1e-8 * np.eye(2)) > 0) for b in r.X)
    ra = O.solve(prob, O.Options())
    assert ra.status == "Optimal" and abs(ra.pobj - float(g["t1_popt"])) < 1e-8


@pytest.mark.parametrize("n", range(4))
def test_schur_brute_force_tiny(n):
    g = golden()
    A, X, Y, S = g[f"tiny{n}_A"], g[f"tiny{n}_X"], g[f"tiny{n}_Y"], g[f"tiny{n}_S"]
    m = A.shape[0]
    prob = O.ingest(np.zeros(m), [A], [np.zeros_like(X)])
    Ss = O.schur_build(prob, [O.kchol(X)[0]], [O.kchol(Y)[0]])
    assert rel_fro_lower(Ss, S) <= 1e-13


def test_schur_two_block_fixture():
    """kkt_regressions.jl:488-565 data, expected S by 50-digit brute force."""
    g = golden()
    prob = O.ingest(np.zeros(4), [g["kkt2_A1"], g["kkt2_A2"]], [np.zeros((2, 2))] * 2)
    X = [g["kkt2_X1"], g["kkt2_X2"]]
    Y = [g["kkt2_Y1"], g["kkt2_Y2"]]
    S = O.schur_build(prob, [O.kchol(b)[0] for b in X], [O.kchol(b)[0] for b in Y])
    assert rel_fro_lower(S, g["kkt2_S"]) <= 1e-14


def test_schur_owner_problem_closed_form():
    A, C, c, S_ref = owner_problem()
    prob = O.ingest(c, A, C)
    I6 = np.eye(6)
    S = O.schur_build(prob, [I6] * 8, [I6] * 8)
    assert np.max(np.abs(np.tril(S - S_ref))) / max(np.max(np.abs(S_ref)), 1.0) < 1e-12


def test_planted_transpose_is_caught(monkeypatch):
    """A plausible slip (M_Y transposed) must fail the brute-force pin."""
    g = golden()
    A, X, Y, S = g["tiny2_A"], g["tiny2_X"], g["tiny2_Y"], g["tiny2_S"]
    prob = O.ingest(np.zeros(A.shape[0]), [A], [np.zeros_like(X)])
    bad = O.schur_build(prob, [O.kchol(X)[0]], [O.kchol(Y)[0].T])
    assert rel_fro_lower(bad, S) > 1e-3


def test_kchol_matches_lapack_and_closed_forms():
    rng = np.random.default_rng(0)
    for k in (1, 2, 3, 17):
        A = random_pd(rng, k)
        Lk, ok = O.kchol(A)
        assert ok
        assert np.allclose(Lk, np.linalg.cholesky(A), rtol=1e-14, atol=1e-14)
    assert not O.kchol(np.array([[1.0, 2.0], [2.0, 1.0]]))[1]
    assert not O.kchol(np.array([[-1.0]]))[1]
    assert not O.kchol(-np.eye(3))[1]


def test_regularization_contract():
    """kkt_regressions.jl:927-1000 on the oracle."""
    m = 12
    for smallest in (1.0, 1e-8, 1e-14):
        _, ok, att = O.factor_dense_kkt(regularization_matrix(m, smallest))
        assert ok and att == 0
    for smallest in (0.0, -1e-14, -1e-8):
        _, ok, att = O.factor_dense_kkt(regularization_matrix(m, smallest))
        assert ok and 0 <= att <= 6
    _, ok, att = O.factor_dense_kkt(regularization_matrix(m, -1e-6))
    assert ok and 1 <= att <= 6
    _, ok, att = O.factor_dense_kkt(regularization_matrix(m, -1.0))
    assert not ok and att == 6


def test_solve_and_refine_identities():
    rng = np.random.default_rng(1)
    m = 40
    M = rng.standard_normal((m, m))
    S = M @ M.T + np.eye(m)
    Lf, ok, _ = O.factor_dense_kkt(S)
    r = rng.standard_normal(m)
    dx = O.solve_dense_kkt(Lf, r)
    assert np.allclose(S @ dx, r, rtol=1e-11, atol=1e-11)
    res, rho = O.direction_residual(S, r, dx)
    assert np.allclose(rho, r - S @ dx)
    dx2, steps, res2 = O.refine_direction(S, Lf, r, dx + 1e-6 * rng.standard_normal(m), O.Options())
    assert steps >= 1 and res2 < 1e-9


def test_fraction_to_boundary_matches_eigen_bound():
    rng = np.random.default_rng(2)
    for k in (3, 8, 20):
        X = random_pd(rng, k)
        dX = rng.standard_normal((k, k))
        dX = 3 * (dX + dX.T)
        Lx = np.linalg.cholesky(X)
        W = sla.solve_triangular(Lx, sla.solve_triangular(Lx, dX, lower=True).T, lower=True)
        lam = np.linalg.eigvalsh((W + W.T) / 2).min()
        exact = min(1.0, -1.0 / lam) if lam < 0 else 1.0
        t = O.fraction_to_boundary_bound(X, dX)
        assert abs(t - exact) <= 1e-10 * max(exact, 1e-300)
    # 2 x 2 closed form vs the determinant root
    X = np.array([[1.0, 0.2], [0.2, 0.5]])
    dX = np.array([[-2.0, 0.3], [0.3, 0.1]])
    t = O.fraction_to_boundary_bound(X, dX)
    assert abs(np.linalg.det(X + t * dX)) < 1e-12 or abs((X + t * dX)[0, 0]) < 1e-12


def test_backtracking_is_first_accepted_power():
    rng = np.random.default_rng(3)
    k = 6
    X = random_pd(rng, k)
    dX = -2.5 * X + 0.1 * rng.standard_normal((k, k))
    dX = (dX + dX.T) / 2
    tX, _ = O.backtracking_line_search(O.Problem(np.zeros(1), [np.eye(k)], [np.zeros((k * k, 1))], [k]),
                                       [X], [X], [dX], [dX], 0.9, 1e-10, 0.0)
    assert O.trial_isposdef(X, tX, dX)
    assert not O.trial_isposdef(X, tX / 0.9, dX)


@pytest.mark.parametrize("k,m", [(20, 60), (50, 200)])
def test_dense_generator_solves_with_certificate(k, m):
    """cfg0 shape: strictly feasible start, optimal with a valid certificate, strong duality."""
    g = G.dense_feasible(k, m, seed=0)
    prob = O.ingest(g.c, [g.A_blocks()], [g.C])
    tol = 1e-8
    r = O.solve(prob, O.Options(eps_gap=tol, eps_primal=tol, eps_dual=tol), x0=g.x0, X0=[g.X0], Y0=[g.Y0])
    assert r.status == "Optimal"
    cert = O.certificate(prob, r, tol, tol, tol)
    assert cert["valid"], cert
    assert abs(r.pobj - r.dobj) <= 1e-7 * max(1.0, abs(r.pobj))
    assert r.iterations < 40


def test_generator_start_is_feasible():
    g = G.dense_feasible(10, 25, seed=4)
    prob = O.ingest(g.c, [g.A_blocks()], [g.C])
    p, d = O.solution_residuals(prob, g.x0, [g.X0], [g.Y0])
    assert p < 1e-12 and d < 1e-12


def test_stagnation_detector():
    s = O.Stagnation(3, 1e-3)
    for merit in (10.0, 10.0, 10.0):
        assert not s.observe(merit, False, 100)
    assert s.observe(10.0, False, 100) and s.reason == "no_progress"
    s = O.Stagnation(3, 1e-3)
    for merit in (1e4, 1e3, 1e2, 10.0, 1.0):
        assert not s.observe(merit, False, 100)


def test_adaptive_policy_cold_start_uses_fallback_values():
    o = O.Options()
    p = O.AdaptivePolicy.from_options(o)
    d = O.Diagnostics(primal_residual=2.0, dual_residual=0.1, relative_gap=0.5, mu=1.0, mu_aff=0.5,
                      affine_primal_step=0.5, affine_dual_step=0.5)
    par = O.select_adaptive_parameters(p, d, [])
    assert par.sigma == o.beta and par.backtracking_factor == o.gamma and not par.fallback
    d.primal_residual = 0.1
    d.iteration = 3
    par = O.select_adaptive_parameters(p, d, [])
    # ratio 0.5, affine 0.5 -> sigma = clamp(0.8*max(0.25, 0.08) + 0.2*0.1)
    assert math.isclose(par.sigma, 0.8 * 0.25 + 0.2 * 0.1)
    assert par.primal_fraction_to_boundary == 19 / 20


@pytest.mark.parametrize("name,rho,mode,fixed,cap,expected", _H.REFINE_CASES)
def test_refine_direction_steps_follow_the_contraction(name, rho, mode, fixed, cap, expected):
    """_refine_native_direction! (kkt.jl:3891-3963) on a factor of S applied to M = S + rho G: the
    error contracts by rho*u per correction, so the stop rules decide with wide margins."""
    S, M, r = _H.refinement_case(rho, mode)
    Lf, ok, _ = O.factor_dense_kkt(S)
    dx0 = O.solve_dense_kkt(Lf, r)
    res0, _ = O.direction_residual(M, r, dx0)
    o = O.Options(refine_policy="fixed" if fixed else "auto", refine_steps=cap, refine_max_steps=cap)
    dx, steps, res = O.refine_direction(M, Lf, r, dx0, o)
    assert steps == expected
    if expected == 0:
        assert np.array_equal(dx, dx0) and res == res0
    else:
        assert res < res0 * (0.71 if mode == "ones" else 0.5 ** expected)
        # the exact solution of M dx = r is approached, not S's
        assert np.linalg.norm(dx - np.linalg.solve(M, r)) < np.linalg.norm(dx0 - np.linalg.solve(M, r))


# ---------------------------------------------------------------- setup: cold start and Ruiz
def test_cold_start_shift_helpers_reference_values():
    """test/cold_start.jl:134-330 (the Float64 instances)."""
    A = np.array([[1.0]])
    assert O.cold_psd_shift(A)[:3] == (True, 0.0, 1.0)
    A = np.array([[-1.0]])
    ok, shift, margin, scale = O.cold_psd_shift(A)
    assert ok and shift > 0 and margin == A[0, 0] and margin > O.SQRT_EPS * max(1.0, scale)
    for entry, expect_zero in (([[2, 1], [1, 2]], True), ([[2, 2], [2, 2]], False), ([[1, -3], [-3, 1]], False)):
        M = np.array(entry, dtype=float)
        ok, shift, margin, scale = O.cold_psd_shift(M)
        assert ok
        if expect_zero:
            assert shift == 0.0 and abs(margin - 1.0) <= O.SQRT_EPS * 8
        else:
            expected = (M[0, 0] + M[1, 1]) / 2 - math.hypot((M[0, 0] - M[1, 1]) / 2, (M[0, 1] + M[1, 0]) / 2)
            assert shift > 0 and abs(margin - expected) <= O.SQRT_EPS * 16 * scale
            assert margin > O.SQRT_EPS * max(1.0, scale)
    definite = np.array([[5, 1, 0, 0, 1, 0], [1, 6, 1, 0, 0, 0], [0, 1, 5, 1, 0, 0], [0, 0, 1, 6, 1, 0],
                         [1, 0, 0, 1, 5, 1], [0, 0, 0, 0, 1, 6]], dtype=float)
    ok, shift, margin, scale = O.cold_psd_shift(definite.copy())
    assert ok and shift == 0.0 and margin > O.SQRT_EPS * max(1.0, scale)
    indefinite = definite.copy()
    indefinite[0, 0] = -10.0
    ok, shift, margin, scale = O.cold_psd_shift(indefinite)
    assert ok and shift > 0 and margin > O.SQRT_EPS * max(1.0, scale)
    np.linalg.cholesky(indefinite)  # positive definite after the shift
    poisoned = definite.copy()
    poisoned[np.triu_indices(6, 1)] = np.nan  # lower-authoritative storage
    assert O.cold_psd_shift(poisoned)[:2] == (True, 0.0)
    nonfinite = np.ones((2, 2))
    nonfinite[1, 1] = np.inf
    assert not O.cold_psd_shift(nonfinite)[0]
    assert O.cold_psd_shift(np.zeros((0, 0))) == (True, 0.0, 0.0, 1.0)
    # centering shifts (cold_start.jl:240-262)
    ok, ps, ds = O.cold_centering_shifts(1.0, 2.0, 4.0)
    assert ok and math.isclose(ps, 0.125) and math.isclose(ds, 0.25)
    for bad in (0.0, -1.0):
        assert O.cold_centering_shifts(1.0, bad, 4.0) == (False, 0.0, 0.0)
    assert not O.cold_centering_shifts(0.0, 2.0, 4.0)[0]
    assert not O.cold_centering_shifts(math.inf, 2.0, 4.0)[0]
    assert not O.cold_centering_shifts(1.0, 2.0, math.inf)[0]
    # identity mass floor (cold_start.jl:296-330)
    tiny = O.SQRT_EPS
    ok, ps, ds = O.cold_identity_mass_shifts(tiny, tiny, 1)
    assert ok and ps > 0 and ds > 0 and tiny + ps >= 1.0 and tiny + ds >= 1.0
    assert O.cold_identity_mass_shifts(2.0, 3.0, 2) == (True, 0.0, 0.0)
    assert O.cold_identity_mass_shifts(1.0, 4.0, 2) == (True, 0.5, 0.0)
    for bad in (math.inf, math.nan):
        assert not O.cold_identity_mass_shifts(bad, 1.0, 1)[0]


def test_cold_start_kkt_solutions():
    """The identity-metric solves project C onto span{A_i}: X = A(x) - C is orthogonal to every A_i,
    and Y = A(v) satisfies <A_i, Y> = c_i; the shifted point is strictly interior."""
    g = G.dense_feasible(9, 20, seed=4)
    prob = O.ingest(g.c, [g.A_blocks()], [g.C])
    o = O.Options()
    eye = [np.eye(9)]
    S = O.schur_build(prob, eye, eye)
    Lf, ok, _ = O.factor_dense_kkt(S)
    x = O.solve_dense_kkt(Lf, O.contract(prob, 0, prob.C[0]))
    v = O.solve_dense_kkt(Lf, prob.c)
    Xp = O.build_P(prob, 0, x) - prob.C[0]
    Yp = O.build_P(prob, 0, v)
    assert np.max(np.abs(O.contract(prob, 0, Xp))) <= 1e-12 * np.max(np.abs(prob.C[0])) * 20
    assert np.max(np.abs(O.contract(prob, 0, Yp) - prob.c)) <= 1e-12 * np.max(np.abs(prob.c)) * 20
    ok, reason, xc, X, Y, mu, rec = O.cold_start(prob, o)
    assert ok and reason == "none"
    assert np.array_equal(xc, x)
    np.linalg.cholesky(X[0])
    np.linalg.cholesky(Y[0])
    assert math.isclose(mu[0], o.beta * float(np.sum(X[0] * Y[0])) / 9)


def _badly_scaled(k=12, m=30, seed=1):
    rng = np.random.default_rng(0)
    g = G.dense_feasible(k, m, seed=seed)
    d = 10.0 ** rng.integers(-3, 4, k)
    A = g.A_blocks() * np.outer(d, d)[None]
    return O.ingest(g.c * 10.0 ** rng.integers(-2, 3, m), [A], [g.C * np.outer(d, d)])


def test_equilibrate_identity_on_balanced_data():
    """Row norms already 1 (entries +-1 on every row of C = I and every A_i): e = 1 on every pass, so
    E = 1 after the two minimum passes and s_i = max|A_i| = 1."""
    k, m = 5, 7
    rng = np.random.default_rng(2)
    A = np.sign(rng.standard_normal((m, k, k)))
    A = np.triu(A) + np.triu(A, 1).transpose(0, 2, 1)
    c = rng.standard_normal(m)
    sp, eq = O.equilibrate(O.ingest(c, [A], [np.eye(k)]))
    assert eq.passes == [2] and np.array_equal(eq.E[0], np.ones(k)) and np.array_equal(eq.s, np.ones(m))
    assert eq.objective_scale == np.max(np.abs(c)) and np.array_equal(sp.c, c / np.max(np.abs(c)))


def test_equilibrate_balances_and_preserves_the_optimum():
    """A badly scaled problem (rows scaled by 1e-3..1e3): after the Ruiz passes every block row has
    infinity-norm near 1, and the equilibrated solve maps back to an optimum of the original problem
    (the unscaled solve stalls on it)."""
    prob = _badly_scaled()
    sp, eq = O.equilibrate(prob)
    T = sp.Av[0].T.reshape(prob.m, 12, 12)
    rn = np.maximum(np.max(np.abs(T), axis=(0, 1)), np.max(np.abs(sp.C[0]), axis=1))
    # rows before the per-variable scale: within 1/20 in log after convergence, then the A_i are divided by s
    assert np.all(rn <= 1.0 + 1e-12) and np.all(rn >= 0.5)
    o = O.Options(eps_gap=1e-9, eps_primal=1e-9, eps_dual=1e-9)
    r, _ = O.solve_equilibrated(prob, o, start="cold")
    assert r.status == "Optimal"
    cert = O.certificate(prob, r, 1e-7, 1e-7, 1e-7)
    assert cert["valid"], cert
    assert O.solve(prob, o).status != "Optimal"


# ---------------------------------------------------------------- equality-constrained KKT (n_eq > 0)
def _eq_case(m=60, n=7, seed=3):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((m, m))
    S = M @ M.T / m + np.eye(m)
    B = rng.standard_normal((m, n)) * 10.0 ** rng.integers(-3, 4, n)[None, :]
    return S, B, rng.standard_normal(m), rng.standard_normal(n)


def test_equality_kkt_solves_the_bordered_system():
    """dx, dy against a direct solve of [S -B; B' 0] [dx; dy] = [r; p] (LAPACK, independent of the
    Schur-complement route), and the residual identities."""
    S, B, r, p = _eq_case()
    Lf, ok, _ = O.factor_dense_kkt(S)
    st, Bh, D, LQ, tau = O.factor_equality(Lf, B, 1e-8)
    assert st == 0
    dx, dy = O.solve_equality(Lf, Bh, D, LQ, r, p)
    m, n = B.shape
    K = np.block([[S, -B], [B.T, np.zeros((n, n))]])
    ref = np.linalg.solve(K, np.concatenate([r, p]))
    assert np.allclose(dx, ref[:m], rtol=1e-9, atol=1e-11) and np.allclose(dy, ref[m:], rtol=1e-9, atol=1e-11)
    assert np.max(np.abs(S @ dx - B @ dy - r)) <= 1e-10 * np.max(np.abs(r))
    assert np.max(np.abs(B.T @ dx - p)) <= 1e-10 * max(1.0, np.max(np.abs(p)))
    # the normalized columns have infinity-norm one; a rescaled equality leaves dx unchanged
    assert np.allclose(np.max(np.abs(Bh), axis=0), 1.0, rtol=1e-15)
    B2, p2 = B.copy(), p.copy()
    B2[:, 2] *= 1e3
    p2[2] *= 1e3
    st2, Bh2, D2, LQ2, _ = O.factor_equality(Lf, B2, 1e-8)
    dx2, dy2 = O.solve_equality(Lf, Bh2, D2, LQ2, r, p2)
    assert np.allclose(dx2, dx, rtol=1e-10, atol=1e-12) and math.isclose(dy2[2] * 1e3, dy[2], rel_tol=1e-9)


def test_equality_rank_gate():
    """A duplicated or proportional equality is numerically rank deficient (status 2); tau follows
    _equality_cholesky_rank_tau (kkt.jl:103-112)."""
    S, B, r, p = _eq_case(n=5)
    Lf, _, _ = O.factor_dense_kkt(S)
    assert math.isclose(O.equality_rank_tau(60, 5, 1e-8), 1e-8)
    assert O.equality_rank_tau(60, 5, 1e-20) == 60 * O.EPS
    assert O.equality_rank_tau(60, 5, 0.0) == O.SQRT_EPS
    dup = B.copy()
    dup[:, 4] = dup[:, 1]
    assert O.factor_equality(Lf, dup, 1e-8)[0] in (1, 2)
    prop = B.copy()
    prop[:, 3] = -2.5 * prop[:, 0]
    assert O.factor_equality(Lf, prop, 1e-8)[0] in (1, 2)
    assert O.factor_equality(Lf, B, 1e-8)[0] == 0


def test_condition_estimates_bounds_and_closed_forms():
    """schur_condition_estimate (dpocon) and triangular_condition_estimate (dtrcon product): the exact
    cond_1 of a diagonal S (max/min ratio) and of L = diag, a lower bound of the exact cond_1(S) within
    the estimator's factor 3 on random S, and cond_1(S) <= cond_1(L) cond_inf(L)."""
    d = np.array([4.0, 1.0, 0.25, 9.0])
    S = np.diag(d)
    L = np.diag(np.sqrt(d))
    assert O.schur_condition_estimate(S, L) == pytest.approx(36.0, rel=1e-14)
    assert O.triangular_condition_estimate(L) == pytest.approx(36.0, rel=1e-14)
    rng = np.random.default_rng(3)
    for n, cond in ((20, 1e2), (150, 1e6)):
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        S = (Q * np.geomspace(1.0, cond, n)) @ Q.T
        Lf = np.linalg.cholesky(S)
        exact = np.linalg.cond(S, 1)
        est = O.schur_condition_estimate(S, Lf)
        assert exact / 3 <= est <= exact * (1 + 1e-10)
        prod = np.linalg.cond(Lf, 1) * np.linalg.cond(Lf, np.inf)
        tr = O.triangular_condition_estimate(Lf)
        assert tr <= prod * (1 + 1e-10) and exact <= prod * (1 + 1e-10)
