"""Implicit-differentiation adjoint: the CPU statement pinned by central finite
differences of the central-path map, and the HIP batched adjoint against it.

There is no adjoint in the reference; parity with it is unpinned.  The pins
are mathematical: dL/dC and dL/db agree with central differences (step 1e-6)
of L at the central-path point of the final mu, to 1e-6 relative.
"""

import numpy as np
import pytest

import adjoint_oracle as AO
import sdpx_oracle as O

from sdpx_mi355x import generators as G


def _unit_sym(rng, k):
    E = rng.standard_normal((k, k))
    E = (E + E.T) / 2
    return E / np.linalg.norm(E)


def _unit(rng, n):
    v = rng.standard_normal(n)
    return v / np.linalg.norm(v)


def _solved_central(g, tol=1e-2):
    """A central-path point: the oracle's iterate at a moderate gap, re-centred at its own mu.

    The finite-difference step (1e-6) must be small against mu: the central-path
    point bends on the scale of mu, so the check runs at mu ~ 1e-3."""
    A = g.A_blocks() if isinstance(g, G.DenseSDP) else g.dense_rows().reshape(g.m, g.k, g.k)
    prob = O.ingest(g.c, [A], [g.C])
    r = O.solve(prob, O.Options(eps_gap=tol, eps_primal=tol, eps_dual=tol), x0=g.x0, X0=[np.eye(g.k)],
                Y0=[np.eye(g.k)])
    assert r.status == "Optimal"
    mu = float(np.sum(r.X[0] * r.Y[0])) / g.k
    x, X, Y = AO.center(prob, r.x, r.X[0], r.Y[0], mu, iterations=6)
    return prob, x, X, Y, mu


def _fd_check(prob, x, X, Y, mu, Gm, h, rng, step=1e-6):
    dC, db = AO.adjoint(prob, x, X, Y, Gm, h)
    E = _unit_sym(rng, X.shape[0])
    e = _unit(rng, prob.m)
    lp = AO.central_point_loss(AO.perturbed(prob, dC_std=step * E), x, X, Y, mu, Gm, h)
    lm = AO.central_point_loss(AO.perturbed(prob, dC_std=-step * E), x, X, Y, mu, Gm, h)
    fd_C = (lp - lm) / (2 * step)
    lp = AO.central_point_loss(AO.perturbed(prob, db=step * e), x, X, Y, mu, Gm, h)
    lm = AO.central_point_loss(AO.perturbed(prob, db=-step * e), x, X, Y, mu, Gm, h)
    fd_b = (lp - lm) / (2 * step)
    an_C = float(np.sum(dC * E))
    an_b = float(np.dot(db, e))
    # directions are unit vectors: errors are relative to the gradient norms
    return (an_C, fd_C, float(np.linalg.norm(dC))), (an_b, fd_b, float(np.linalg.norm(db)))


def _rel(a, b, scale):
    return abs(a - b) / max(scale, 1e-300)


@pytest.mark.parametrize("k,m,seed", [(6, 10, 0), (12, 30, 1)])
def test_oracle_adjoint_matches_central_differences(k, m, seed):
    g = G.dense_feasible(k, m, seed=seed)
    prob, x, X, Y, mu = _solved_central(g)
    rng = np.random.default_rng(seed)
    Gm, h = _unit_sym(rng, k), _unit(rng, m)
    (aC, fC, nC), (ab, fb, nb) = _fd_check(prob, x, X, Y, mu, Gm, h, rng)
    assert _rel(aC, fC, nC) <= 1e-6, (aC, fC, nC)
    assert _rel(ab, fb, nb) <= 1e-6, (ab, fb, nb)


def test_planted_adjoint_sign_error_is_caught(monkeypatch):
    g = G.dense_feasible(6, 10, seed=0)
    prob, x, X, Y, mu = _solved_central(g)
    rng = np.random.default_rng(0)
    Gm, h = _unit_sym(rng, 6), _unit(rng, 10)
    orig = AO.adjoint
    monkeypatch.setattr(AO, "adjoint", lambda *a: (lambda r: (r[0], -r[1]))(orig(*a)))
    _, (ab, fb, nb) = _fd_check(prob, x, X, Y, mu, Gm, h, rng)
    assert _rel(ab, fb, nb) > 1e-3


@pytest.mark.gpu
def test_gpu_adjoint_matches_oracle_and_central_differences(gpu_lib):
    """The differentiable-layer shape: 16 problems n=32, m=128 (C, b perturbed), gap 1e-9."""
    from sdpx_mi355x.batch import BatchOptions, BatchSolver
    from sdpx_mi355x.policy import SolverOptions

    probs = G.batched_dense_layer(16, seed=11)
    B = len(probs)
    rng = np.random.default_rng(5)
    Gs = [_unit_sym(rng, 32) for _ in range(B)]
    hs = [_unit(rng, 128) for _ in range(B)]
    tol = 1e-9
    bo = BatchOptions.from_solver_options(SolverOptions(eps_gap=tol, eps_primal=tol, eps_dual=tol))
    bo.adjoint_mu = 1e-2  # central differences with step 1e-6 are accurate to 1e-6 only for mu >> 1e-6
    with BatchSolver([32] * B, [128] * B) as bs:
        bs.set_dense([g.Av_rows for g in probs], [g.C for g in probs], [g.c for g in probs])
        bs.set_start([g.x0 for g in probs], [np.eye(32)] * B, [np.eye(32)] * B)
        res = bs.solve(bo)
        restored, ck_mu = bs.restore_checkpoint()
        assert restored >= 12
        bs.center(4)
        cen = bs.result()
        dC, db = bs.adjoint(Gs, hs)
    checked = 0
    for p, g in enumerate(probs):
        if res.status[p] != "Optimal":
            continue
        prob = O.ingest(g.c, [g.A_blocks()], [g.C])
        x, X, Y = cen.x[p], cen.X[p], cen.Y[p]
        # same formula on the CPU at the GPU's iterate
        rC, rb = AO.adjoint(prob, x, X, Y, Gs[p], hs[p])
        eC = np.linalg.norm(dC[p] - rC) / np.linalg.norm(rC)
        eb = np.linalg.norm(db[p] - rb) / np.linalg.norm(rb)
        # finite differences of the central-path map at the GPU's mu
        mu = float(np.sum(X * Y)) / 32
        xs, Xs, Ys = AO.center(prob, x, X, Y, mu, iterations=3)
        (aC, fC, nC), (ab, fb, nb) = _fd_check(prob, xs, Xs, Ys, mu, Gs[p], hs[p], np.random.default_rng(100 + p))
        print(f"problem {p}: mu {mu:.1e} |dC - cpu|/|cpu| = {eC:.2e}, |db - cpu|/|cpu| = {eb:.2e}; FD rel C "
              f"{_rel(aC, fC, nC):.2e} b {_rel(ab, fb, nb):.2e}", flush=True)
        assert eC <= 1e-8 and eb <= 1e-8
        assert _rel(aC, fC, nC) <= 1e-6 and _rel(ab, fb, nb) <= 1e-6
        checked += 1
    assert checked >= 12


@pytest.mark.parametrize("p", [1, 4, 8])
def test_central_path_gradient_approaches_the_optimum(p):
    """The adjoint is the exact derivative of the central-path point of mu (pinned above by central
    differences); the optimum's derivative is its mu -> 0 limit.  On the differentiable-layer shape
    (n=32, m=128, C and b perturbed) the oracle stalls near gap 1e-11 in FP64, so central differences
    of its optimum map with step 1e-6 carry the solve error 1e-11 / 1e-6, and the limit is measured
    instead: dL/db changes 10x less per decade of mu (linear convergence, ~1e-4 relative left at
    mu = 1e-6), and dL/dC changes ~1e-3 between mu = 1e-5 and 1e-6, below which FP64 cancellation
    (lambda's error, kappa(S) eps, amplified by 1/mu in X^-1 A_lambda Y) sets its floor."""
    g = G.batched_dense_layer(16, seed=11)[p]
    prob = O.ingest(g.c, [g.A_blocks()], [g.C])
    rng = np.random.default_rng(5 + p)
    Gm, h = _unit_sym(rng, 32), _unit(rng, 128)
    r = O.solve(prob, O.Options(eps_gap=1e-9, eps_primal=1e-9, eps_dual=1e-9), x0=g.x0, X0=[np.eye(32)],
                Y0=[np.eye(32)])
    assert r.status == "Optimal"
    D = {}
    for mu in (1e-4, 1e-5, 1e-6, 1e-7):
        x, X, Y = AO.center(prob, r.x, r.X[0], r.Y[0], mu, iterations=12)
        assert abs(float(np.sum(X * Y)) / 32 - mu) <= 1e-9 * mu
        D[mu] = AO.adjoint(prob, x, X, Y, Gm, h)

    def change(a, b, i):
        return float(np.linalg.norm(D[a][i] - D[b][i]) / np.linalg.norm(D[b][i]))

    b_changes = [change(1e-4, 1e-5, 1), change(1e-5, 1e-6, 1), change(1e-6, 1e-7, 1)]
    c_changes = [change(1e-4, 1e-5, 0), change(1e-5, 1e-6, 0), change(1e-6, 1e-7, 0)]
    print(f"problem {p}: db change per decade {b_changes}, dC change per decade {c_changes}")
    assert 0.05 <= b_changes[1] / b_changes[0] <= 0.2 and 0.05 <= b_changes[2] / b_changes[1] <= 0.2
    assert b_changes[2] <= 1e-3
    assert c_changes[1] <= 3e-3 and c_changes[1] < c_changes[0]
