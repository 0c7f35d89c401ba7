"""Mixed precision (FP32 or BF16x3 Schur Gram + FP64 matrix-free refinement) on the dense route.

Pins: the matrix-free FP64 residual equals the assembled-S residual when S is
assembled in FP64; the FP32-assembled S is within FP32 accuracy of the oracle's
S and the BF16x3-assembled S within the 2^-16 accuracy of its hi / lo split
(a dropped cross term or plane would leave 2^-8); mixed solves reach the same
optimum as the oracle (1e-8) with a valid certificate and at least one
reduced-precision iteration.
"""

import numpy as np
import pytest

import sdpx_oracle as O

from sdpx_mi355x import MI355XDenseBackend, SolverOptions, generators, ingest, solve

from _helpers import rel_fro_lower


def _state(k, m, seed):
    g = generators.dense_feasible(k, m, seed=seed)
    prob = ingest(g.c, [g.A_blocks()], [g.C])
    oprob = O.ingest(g.c, [g.A_blocks()], [g.C])
    rng = np.random.default_rng(seed)
    Q = rng.standard_normal((k, k))
    X = Q @ Q.T / k + 0.5 * np.eye(k)
    Q = rng.standard_normal((k, k))
    Y = Q @ Q.T / k + 0.5 * np.eye(k)
    return g, prob, oprob, X, Y


@pytest.mark.gpu
def test_exact_residual_matches_assembled_operator(gpu_lib):
    g, prob, oprob, X, Y = _state(40, 150, 1)
    b = MI355XDenseBackend(prob)
    try:
        b.set_iterate(g.x0, [X], [Y])
        mu = np.array([0.1 * float(np.sum(X * Y)) / 40])
        b.compute_residuals(mu, affine=False, factor=True)
        b.schur_build()
        assert b.factorize()[0]
        b.predictor_rhs()
        b.solve_direction()
        # perturb dx so the residual is not at rounding level
        dx, _, _ = b.get_direction()
        b.set_dx(dx * (1.0 + 1e-3))
        res_s, rn = b.direction_residual()
        res_e, rn2 = b.exact_residual()
        assert rn == rn2
        assert abs(res_s - res_e) <= 1e-10 * max(res_s, 1e-300) + 1e-13 * max(rn, 1.0)
    finally:
        b.close()


@pytest.mark.gpu
def test_fp32_schur_within_fp32_accuracy(gpu_lib):
    g, prob, oprob, X, Y = _state(64, 300, 2)
    b = MI355XDenseBackend(prob)
    try:
        b.set_iterate(g.x0, [X], [Y])
        b.compute_residuals(np.array([0.0]), affine=True, factor=True)
        b.set_precision(1)
        b.schur_build()
        S32 = b.get_schur()
        b.set_precision(0)
        b.schur_build()
        S64 = b.get_schur()
    finally:
        b.close()
    LX, _ = O.kchol(X)
    MY, _ = O.kchol(Y)
    ref = O.schur_build(oprob, [LX], [MY])
    assert rel_fro_lower(S64, ref) <= 1e-12
    err = rel_fro_lower(S32, ref)
    print(f"FP32-assembled Schur rel Frobenius error {err:.2e}", flush=True)
    assert 1e-10 < err <= 2e-5  # genuinely FP32, and within FP32 accuracy


@pytest.mark.gpu
@pytest.mark.parametrize("factor32", [False, True])
@pytest.mark.parametrize("mode", ["auto", "on"])
@pytest.mark.parametrize("k,m", [(50, 200), (64, 500)])
def test_mixed_solve_matches_oracle(gpu_lib, k, m, mode, factor32):
    """mode "auto": the condition estimate gates each reduced-precision step; "on": the measured
    trial, where the residual and refinement guards alone decide.  factor32: the reduced-precision
    steps also factor S in FP32."""
    g = generators.dense_feasible(k, m, seed=3)
    prob = ingest(g.c, [g.A_blocks()], [g.C])
    tol = 1e-8
    o = SolverOptions(eps_gap=tol, eps_primal=tol, eps_dual=tol, precision="mixed", mixed_mode=mode,
                      mixed_factor=factor32)
    res = solve(prob, o, x0=g.x0, X0=[g.X0], Y0=[g.Y0])
    oprob = O.ingest(g.c, [g.A_blocks()], [g.C])
    ref = O.solve(oprob, O.Options(eps_gap=tol, eps_primal=tol, eps_dual=tol), x0=g.x0, X0=[g.X0], Y0=[g.Y0])
    print(f"mixed: {res.status} it={res.iterations} fp32={res.fp32_iterations} fp64={res.fp64_iterations} "
          f"pobj={res.pobj:.12e} | oracle {ref.status} it={ref.iterations} pobj={ref.pobj:.12e}", flush=True)
    assert res.status == ref.status == "Optimal"
    assert res.fp32_iterations >= 1
    assert abs(res.pobj - ref.pobj) <= 1e-8 * max(1.0, abs(ref.pobj))
    assert abs(res.iterations - ref.iterations) <= 3
    cert = O.certificate(oprob, O.Result(res.status, res.x, res.X, res.Y, res.pobj, res.dobj, res.iterations,
                                         res.p_res, res.d_res, res.gap_rel), tol, tol, tol)
    assert cert["valid"], cert


@pytest.mark.gpu
@pytest.mark.parametrize("k,m", [(40, 130), (30, 300), (20, 64)])
def test_fp32_factor_solves_to_fp32_accuracy(gpu_lib, k, m):
    """The FP32 Schur factor (FP32 MFMA panels and trailing updates, FP64 diagonal blocks): the
    solve's backward error sits at FP32 level (it is not the FP64 factor), refinement against the
    FP64 operator recovers 1e-10, and the condition estimate from the FP32 factor is LAPACK's
    dpocon estimate of cond_1(S) from the FP64 factor to within FP32 rounding."""
    from scipy.linalg import lapack

    g, prob, oprob, X, Y = _state(k, m, 11)
    LX, _ = O.kchol(X)
    MY, _ = O.kchol(Y)
    S = O.schur_build(oprob, [LX], [MY])
    S = np.tril(S) + np.tril(S, -1).T
    r = np.random.default_rng(3).standard_normal(m)
    b = MI355XDenseBackend(prob)
    try:
        b.set_iterate(g.x0, [X], [Y])
        b.compute_residuals(np.array([0.0]), affine=True, factor=True)
        b.set_schur(S)
        out = {}
        for prec in (0, 1):
            b.set_factor_precision(prec)
            ok, attempts = b.factorize()
            assert ok and attempts == 0
            b.set_rhs(r)
            b.solve_direction()
            dx = b.get_direction()[0]
            out[prec] = (np.linalg.norm(S @ dx - r) / (np.linalg.norm(S) * np.linalg.norm(dx)),
                         b.factor_condition(1)[0])
        b.set_factor_precision(1)
        b.factorize()
        b.set_rhs(r)
        b.solve_direction()
        steps, res = b.refine_exact(1e-10 * max(np.max(np.abs(r)), 1.0), 32)
    finally:
        b.close()
    Lc, info = lapack.dpotrf(S, lower=1)
    assert info == 0
    rcond, info = lapack.dpocon(Lc, np.max(np.sum(np.abs(S), axis=0)), uplo="L")
    print(f"k={k} m={m}: backward error fp64 {out[0][0]:.2e} fp32 {out[1][0]:.2e}; cond est fp64 "
          f"{out[0][1]:.4e} fp32 {out[1][1]:.4e} lapack {1 / rcond:.4e}; refined in {steps} to {res:.2e}", flush=True)
    assert out[0][0] <= 1e-14
    assert 1e-12 < out[1][0] <= 1e-5
    assert abs(out[1][1] - 1 / rcond) <= 1e-3 * (1 / rcond)
    assert res <= 1e-10 * max(np.max(np.abs(r)), 1.0)


@pytest.mark.gpu
def test_bf16x3_schur_within_split_accuracy(gpu_lib):
    g, prob, oprob, X, Y = _state(64, 300, 2)
    b = MI355XDenseBackend(prob)
    try:
        b.set_iterate(g.x0, [X], [Y])
        b.compute_residuals(np.array([0.0]), affine=True, factor=True)
        b.set_precision(2)
        b.schur_build()
        Sb = b.get_schur()
        b.set_precision(1)
        b.schur_build()
        S32 = b.get_schur()
    finally:
        b.close()
    LX, _ = O.kchol(X)
    MY, _ = O.kchol(Y)
    ref = O.schur_build(oprob, [LX], [MY])
    err = rel_fro_lower(Sb, ref)
    err32 = rel_fro_lower(S32, ref)
    print(f"BF16x3-assembled Schur rel Frobenius error {err:.2e} (FP32 {err32:.2e})", flush=True)
    assert err32 < err <= 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("k", [64, 50])
def test_bf16x3_fp32_panels_within_split_accuracy(gpu_lib, k):
    """P-hat on FP32 MFMA (bf16 planes written by the product): the BF16x3-assembled S stays
    within the split's accuracy of the oracle's; before and after equilibration (the FP32 copy
    of Av follows the rescaled panel).  k = 50 is not a multiple of 4 and keeps the FP64
    products: its S equals the FP64-product one bit for bit."""
    g, prob, oprob, X, Y = _state(k, 300, 2)
    b = MI355XDenseBackend(prob)
    try:
        for equilibrated in (False, True):
            if equilibrated:
                b.equilibrate(2)
                oprob, _ = O.equilibrate(oprob, 2)  # bitwise the device's scaling
            b.set_iterate(g.x0, [X], [Y])
            b.compute_residuals(np.array([0.0]), affine=True, factor=True)
            b.set_precision(2)
            b.set_panel_precision(0)
            b.schur_build()
            S64p = b.get_schur()
            b.set_panel_precision(1)
            b.schur_build()
            S32p = b.get_schur()
            LX, _ = O.kchol(X)
            MY, _ = O.kchol(Y)
            ref = O.schur_build(oprob, [LX], [MY])
            e64, e32 = rel_fro_lower(S64p, ref), rel_fro_lower(S32p, ref)
            print(f"k={k} equilibrated={equilibrated}: BF16x3 Schur rel error, FP64 products {e64:.2e}, "
                  f"FP32 products {e32:.2e}", flush=True)
            if k % 4:
                assert np.array_equal(np.tril(S32p), np.tril(S64p))
            assert e32 <= 2.0 * e64 and e32 <= 1e-4
    finally:
        b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("k2nd", [48, 50])
def test_bf16x3_fp32_panels_two_blocks(gpu_lib, k2nd):
    """Two dense blocks (64 and 48 or 50) share the bf16 planes: the second block's shorter
    panels leave no stale rows, and with k = 50 one block takes the FP32 products and the
    other the FP64 ones within the same Schur matrix."""
    m = 180
    blocks = [(64, 5), (k2nd, 6)]
    gs = [generators.dense_feasible(k, m, seed=sd) for k, sd in blocks]
    c = sum(g.c for g in gs)
    A = [g.A_blocks() for g in gs]
    Cs = [g.C for g in gs]
    prob = ingest(c, A, Cs)
    oprob = O.ingest(c, A, Cs)
    rng = np.random.default_rng(7)
    Xs, Ys = [], []
    for k, _ in blocks:
        Q = rng.standard_normal((k, k))
        Xs.append(Q @ Q.T / k + 0.5 * np.eye(k))
        Q = rng.standard_normal((k, k))
        Ys.append(Q @ Q.T / k + 0.5 * np.eye(k))
    b = MI355XDenseBackend(prob)
    try:
        b.set_iterate(np.zeros(m), Xs, Ys)
        b.compute_residuals(np.array([0.0, 0.0]), affine=True, factor=True)
        b.set_precision(2)
        b.set_panel_precision(0)
        b.schur_build()
        S64p = b.get_schur()
        b.set_panel_precision(1)
        b.schur_build()
        S32p = b.get_schur()
    finally:
        b.close()
    LX = [O.kchol(X)[0] for X in Xs]
    MY = [O.kchol(Y)[0] for Y in Ys]
    ref = O.schur_build(oprob, LX, MY)
    e64, e32 = rel_fro_lower(S64p, ref), rel_fro_lower(S32p, ref)
    print(f"blocks 64 + {k2nd}: BF16x3 Schur rel error, FP64 products {e64:.2e}, FP32 products {e32:.2e}", flush=True)
    assert e32 <= 2.0 * e64 and e32 <= 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["auto", "on"])
def test_mixed_bf16x3_fp32_panels_solve_matches_oracle(gpu_lib, mode):
    k, m = 48, 200
    g = generators.dense_feasible(k, m, seed=3)
    prob = ingest(g.c, [g.A_blocks()], [g.C])
    tol = 1e-8
    o = SolverOptions(eps_gap=tol, eps_primal=tol, eps_dual=tol, precision="mixed_bf16x3", mixed_mode=mode,
                      mixed_panels32=True)
    res = solve(prob, o, x0=g.x0, X0=[g.X0], Y0=[g.Y0])
    oprob = O.ingest(g.c, [g.A_blocks()], [g.C])
    ref = O.solve(oprob, O.Options(eps_gap=tol, eps_primal=tol, eps_dual=tol), x0=g.x0, X0=[g.X0], Y0=[g.Y0])
    print(f"mixed_bf16x3 (FP32 products): {res.status} it={res.iterations} low={res.fp32_iterations} "
          f"fp64={res.fp64_iterations} pobj={res.pobj:.12e} | oracle {ref.status} it={ref.iterations} "
          f"pobj={ref.pobj:.12e}", flush=True)
    assert res.status == ref.status == "Optimal"
    assert res.fp32_iterations >= 1
    assert abs(res.pobj - ref.pobj) <= 1e-8 * max(1.0, abs(ref.pobj))
    assert abs(res.iterations - ref.iterations) <= 3
    cert = O.certificate(oprob, O.Result(res.status, res.x, res.X, res.Y, res.pobj, res.dobj, res.iterations,
                                         res.p_res, res.d_res, res.gap_rel), tol, tol, tol)
    assert cert["valid"], cert


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["auto", "on"])
def test_mixed_bf16x3_solve_matches_oracle(gpu_lib, mode):
    k, m = 50, 200
    g = generators.dense_feasible(k, m, seed=3)
    prob = ingest(g.c, [g.A_blocks()], [g.C])
    tol = 1e-8
    o = SolverOptions(eps_gap=tol, eps_primal=tol, eps_dual=tol, precision="mixed_bf16x3", mixed_mode=mode,
                      mixed_panels32=False)
    res = solve(prob, o, x0=g.x0, X0=[g.X0], Y0=[g.Y0])
    oprob = O.ingest(g.c, [g.A_blocks()], [g.C])
    ref = O.solve(oprob, O.Options(eps_gap=tol, eps_primal=tol, eps_dual=tol), x0=g.x0, X0=[g.X0], Y0=[g.Y0])
    print(f"mixed_bf16x3: {res.status} it={res.iterations} low={res.fp32_iterations} fp64={res.fp64_iterations} "
          f"pobj={res.pobj:.12e} | oracle {ref.status} it={ref.iterations} pobj={ref.pobj:.12e}", flush=True)
    assert res.status == ref.status == "Optimal"
    assert res.fp32_iterations >= 1
    assert abs(res.pobj - ref.pobj) <= 1e-8 * max(1.0, abs(ref.pobj))
    assert abs(res.iterations - ref.iterations) <= 3
    cert = O.certificate(oprob, O.Result(res.status, res.x, res.X, res.Y, res.pobj, res.dobj, res.iterations,
                                         res.p_res, res.d_res, res.gap_rel), tol, tol, tol)
    assert cert["valid"], cert


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["random", "graded"])
def test_condition_estimates_match_lapack(gpu_lib, case):
    """sdpx_dense_factor_condition against LAPACK on the oracle's factor of the same S: dpocon's
    cond_1(S) and the dtrcon product of _triangular_condition_estimate (mixed_precision_kkt.jl:371-392).
    Higham's estimator takes the same steps as LAPACK's, so the two agree unless a rounding-level tie
    changes a step: measured equal to 1e-12; the bar is 1e-8."""
    m = 300
    rng = np.random.default_rng(7)
    Q, _ = np.linalg.qr(rng.standard_normal((m, m)))
    ev = np.exp(rng.uniform(0, np.log(1e3 if case == "random" else 1e9), m))
    S = (Q * ev) @ Q.T
    _, prob, _, X, Y = _state(8, m, 2)
    b = MI355XDenseBackend(prob)
    try:
        b.set_schur(S)
        assert b.factorize() == (True, 0)
        cs, ct, r1, ri = b.factor_condition()
    finally:
        b.close()
    Lf, ok, _ = O.factor_dense_kkt(S)
    cs_o, ct_o = O.schur_condition_estimate(S, Lf), O.triangular_condition_estimate(Lf)
    print(f"{case}: cond_1(S) gpu {cs:.6e} lapack {cs_o:.6e} exact {np.linalg.cond(S, 1):.6e}; "
          f"trcon gpu {ct:.6e} lapack {ct_o:.6e}", flush=True)
    assert abs(cs - cs_o) <= 1e-8 * cs_o and abs(ct - ct_o) <= 1e-8 * ct_o
    assert cs <= ct * (1 + 1e-12)
