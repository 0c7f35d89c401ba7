"""The dense route at the benchmark's own size (one PSD block n = 400, m = 4000).

The Schur matrix of the first three Newton steps is compared with the oracle's
at the GPU's own iterate, and the solve with the oracle's solve of the same
problem (a committed fixture).  Properties that do not need the oracle pin the
rest: the assembled FP64 Schur
matrix against the matrix-free FP64 operator Av' vec(X^-1 (sum dx_i A_i) Y)
(two independent GPU computations of S dx), bitwise repeatability of the
assembly, agreement between different tilings (another split-K, tile order
and panel tiles only reorder the sums: a planted index or transpose error
would not stay within 1e-13), a positive definite
factor with no regularisation, and a full solve to rel gap 1e-8 whose
certificate the oracle checks on the CPU.
"""

import json
from pathlib import Path

import numpy as np
import pytest

import sdpx_oracle as O
from sdpx_mi355x import MI355XDenseBackend, SolverOptions, generators, ingest, solve

from _helpers import rel_fro_lower

pytestmark = pytest.mark.gpu

K, M = 400, 4000
FIXTURES = Path(__file__).resolve().parent / "golden" / "oracle_solves.json"


@pytest.fixture(scope="module")
def case():
    g = generators.dense_feasible(K, M, seed=0)
    A = g.A_blocks()
    return g, ingest(g.c, [A], [g.C]), A


def test_fullsize_schur_properties(gpu_lib, case):
    g, prob, _ = case
    rng = np.random.default_rng(1)
    Q = rng.standard_normal((K, K))
    X = Q @ Q.T / K + 0.5 * np.eye(K)
    Q = rng.standard_normal((K, K))
    Y = Q @ Q.T / K + 0.5 * np.eye(K)
    b = MI355XDenseBackend(prob)
    try:
        default_split = b.schur_split()
        b.set_iterate(g.x0, [X], [Y])
        mu = np.array([0.1 * float(np.sum(X * Y)) / K])
        _, _, ok = b.compute_residuals(mu, affine=False, factor=True)
        assert ok
        b.schur_build()
        S0 = b.get_schur()
        b.schur_build()
        assert np.array_equal(np.tril(b.get_schur()), np.tril(S0))  # fixed-order split-K reduction
        # another tiling (split-K 8, 4 x 4 tile order, other panel tiles) reorders every sum
        b.set_tiling(("64x64", "64x128"), split=max(1, default_split // 2), tile_group=4)
        b.schur_build()
        L0, L1 = np.tril(S0), np.tril(b.get_schur())
        rel = np.linalg.norm(L1 - L0) / np.linalg.norm(L0)
        print(f"split {default_split} vs {max(1, default_split // 2)}, other tiles: {rel:.2e}", flush=True)
        assert 0.0 < rel <= 1e-13
        t = b.autotune()
        print(t, flush=True)
        # the autotuner overwrote the panel and Schur buffers: rebuild from the same iterate
        b.set_iterate(g.x0, [X], [Y])
        b.compute_residuals(mu, affine=False, factor=True)
        b.schur_build()
        S1 = b.get_schur()
        rel = np.linalg.norm(np.tril(S1) - L0) / np.linalg.norm(L0)
        print(f"tuned {t['split']}/{t['tile_group']}/{t['panel_tiles']}: {rel:.2e}")
        assert rel <= 1e-13
        assert np.all(np.diag(S1) > 0)
        ok, attempts = b.factorize()
        assert ok and attempts == 0
        b.predictor_rhs()
        b.solve_direction()
        dx, _, _ = b.get_direction()
        b.set_dx(dx * (1.0 + 1e-3))  # off the solution, so the residual is not at rounding level
        res_s, rn = b.direction_residual()
        res_e, rn2 = b.exact_residual()
        print(f"assembled {res_s:.6e} matrix-free {res_e:.6e} (|r| {rn:.3e})")
        assert rn == rn2
        assert abs(res_s - res_e) <= 1e-10 * res_s + 1e-13 * max(rn, 1.0)
    finally:
        b.close()


def test_fullsize_solve_matches_oracle(gpu_lib, case):
    """The full-size FP64 solve against the oracle: the Schur matrix of its first three Newton steps
    against the oracle's schur_build at the GPU's own iterate (<= 1e-12 relative Frobenius), and the
    oracle's own solve of the same problem (tests/golden/oracle_solves.json, written by
    make_solve_fixtures.py): same status, iterations +-2, objective 1e-8 relative, valid certificate."""
    g, prob, A = case
    tol = 1e-8
    o = SolverOptions(eps_gap=tol, eps_primal=tol, eps_dual=tol)
    b = MI355XDenseBackend(prob)
    try:
        b.autotune()
        res = solve(prob, o, x0=g.x0, X0=[g.X0], Y0=[g.Y0], backend=b, trace_schur=3)
    finally:
        b.close()
    print(f"{res.status} iters={res.iterations} pobj={res.pobj:.12e} gap={res.gap_rel:.2e}", flush=True)
    ref = json.loads(FIXTURES.read_text())["dense_n400_m4000"]
    print(f"oracle {ref['status']} iters={ref['iterations']} pobj={ref['pobj']:.12e}", flush=True)
    assert res.status == ref["status"] == "Optimal" and res.gap_rel <= tol
    assert abs(res.iterations - ref["iterations"]) <= 2
    assert abs(res.pobj - ref["pobj"]) <= 1e-8 * max(1.0, abs(ref["pobj"]))
    oprob = O.ingest(g.c, [A], [g.C])
    cert = O.certificate(oprob, O.Result(res.status, res.x, res.X, res.Y, res.pobj, res.dobj, res.iterations,
                                         res.p_res, res.d_res, res.gap_rel), tol, tol, tol)
    assert cert["valid"], cert
    assert len(res.schur_trace) == 3
    for it, t in enumerate(res.schur_trace, 1):
        LX, ok1 = O.kchol(t["X"][0])
        MY, ok2 = O.kchol(t["Y"][0])
        assert ok1 and ok2
        S_ref = O.schur_build(oprob, [LX], [MY])
        rel = rel_fro_lower(t["S"], S_ref)
        print(f"iteration {it}: Schur rel Frobenius vs oracle {rel:.2e}", flush=True)
        assert rel <= 1e-12
