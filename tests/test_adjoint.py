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


def _t1():
    A = np.zeros((2, 2, 2))
    A[0, 0, 0] = A[1, 1, 1] = 1.0
    return O.ingest([2.0, 3.0], [A], [np.array([[0.0, 1.0], [1.0, 0.0]])])


def test_optimum_adjoint_t1_closed_form():
    """The optimum of T1 (test/correctness.jl:71-85) in closed form: x* = (sqrt(c2/c1), sqrt(c1/c2)) times
    C12, x*_1 += C11, x*_2 += C22, Y* = [[c1, sqrt(c1 c2)], [., c2]] (complementary slackness); the
    derivative of L = <G, Y*> + h'x* from these formulas against optimum_adjoint (the implicit function
    theorem on the optimality conditions), to the accuracy of the solve (gap 1e-9)."""
    prob = _t1()
    r = O.solve(prob, O.Options(eps_gap=1e-9, eps_primal=1e-9, eps_dual=1e-9))
    assert r.status == "Optimal"
    rng = np.random.default_rng(0)
    Gm, h = _unit_sym(rng, 2), _unit(rng, 2)
    c1, c2 = 2.0, 3.0
    x1, x2 = np.sqrt(c2 / c1), np.sqrt(c1 / c2)
    dL_dc = np.array([Gm[0, 0] + Gm[0, 1] * np.sqrt(c2 / c1) - h[0] * x1 / (2 * c1) + h[1] * x2 / (2 * c1),
                      Gm[1, 1] + Gm[0, 1] * np.sqrt(c1 / c2) + h[0] * x1 / (2 * c2) - h[1] * x2 / (2 * c2)])
    hx = h[0] * x1 + h[1] * x2
    dL_dC = np.array([[h[0], hx / 2], [hx / 2, h[1]]])
    dC, db = AO.optimum_adjoint(prob, r.x, r.X[0], r.Y[0], Gm, h)
    # the user's model: dL/dC_std = -dL/dC (core), dL/db = -dL/dc (core)
    assert np.max(np.abs(db + dL_dc)) <= 1e-6 and np.max(np.abs(dC + dL_dC)) <= 1e-6
    # a planted slip (Y's block on the null space of X transposed into X's range) is caught
    bad = AO.optimum_adjoint(prob, r.x, r.Y[0], r.X[0], Gm, h)
    assert bad is None or np.max(np.abs(bad[1] + dL_dc)) > 1e-3


def test_central_path_adjoint_converges_linearly_to_the_optimum():
    """On T1 the central-path adjoint approaches optimum_adjoint linearly in mu (one decade of mu, one
    decade of distance), and one Richardson step removes the linear term."""
    prob = _t1()
    r = O.solve(prob, O.Options(eps_gap=1e-9, eps_primal=1e-9, eps_dual=1e-9))
    rng = np.random.default_rng(1)
    Gm, h = _unit_sym(rng, 2), _unit(rng, 2)
    ref = AO.optimum_adjoint(prob, r.x, r.X[0], r.Y[0], Gm, h)
    dist = []
    for mu in (1e-2, 1e-3):
        x, X, Y = AO.center(prob, r.x, r.X[0], r.Y[0], mu, iterations=20)
        D = AO.adjoint(prob, x, X, Y, Gm, h)
        dist.append(np.linalg.norm(D[0] - ref[0]) / np.linalg.norm(ref[0]))
    assert 7 <= dist[0] / dist[1] <= 13
    ex = AO.extrapolated_adjoint(prob, r.x, r.X[0], r.Y[0], Gm, h, 1e-2)
    assert np.linalg.norm(ex[0] - ref[0]) / np.linalg.norm(ref[0]) <= 1e-2 * dist[1]


@pytest.mark.gpu
def test_gpu_adjoint_limit_matches_the_optimum(gpu_lib):
    """The differentiable-layer shape (16 problems n=32, m=128, gap 1e-9): the device's adjoint_limit
    (central-path adjoints at the checkpoint's mu ~ 1e-5 and at a tenth of it, one Richardson step)
    against optimum_adjoint at the device's solution.  Measured ~1e-5 relative; the plain adjoint at the
    checkpoint sits ~1e-3 away."""
    from sdpx_mi355x.batch import BatchOptions, BatchSolver
    from sdpx_mi355x.policy import SolverOptions

    probs = G.batched_dense_layer(16, seed=11)
    B = len(probs)
    rng = np.random.default_rng(5)
    Gs = [_unit_sym(rng, 32) for _ in range(B)]
    hs = [_unit(rng, 128) for _ in range(B)]
    tol = 1e-9
    bo = BatchOptions.from_solver_options(SolverOptions(eps_gap=tol, eps_primal=tol, eps_dual=tol))
    bo.adjoint_mu = 1e-5
    with BatchSolver([32] * B, [128] * B) as bs:
        bs.set_dense([g.Av_rows for g in probs], [g.C for g in probs], [g.c for g in probs])
        bs.set_start([g.x0 for g in probs], [np.eye(32)] * B, [np.eye(32)] * B)
        res = bs.solve(bo)
        bs.restore_checkpoint()
        bs.center(6)
        dCp, dbp = bs.adjoint(Gs, hs)
        dC, db, mu = bs.adjoint_limit(Gs, hs)
    errs, plain = [], []
    for p, g in enumerate(probs):
        if res.status[p] != "Optimal" or np.isnan(mu[p]):
            continue
        prob = O.ingest(g.c, [g.A_blocks()], [g.C])
        ref = AO.optimum_adjoint(prob, res.x[p], res.X[p], res.Y[p], Gs[p], hs[p])
        if ref is None:
            continue
        eC = np.linalg.norm(dC[p] - ref[0]) / np.linalg.norm(ref[0])
        eb = np.linalg.norm(db[p] - ref[1]) / np.linalg.norm(ref[1])
        pC = np.linalg.norm(dCp[p] - ref[0]) / np.linalg.norm(ref[0])
        # strict-complementarity margin: the smallest nonzero eigenvalue of X* and of Y*
        ex, ey = np.linalg.eigvalsh(res.X[p]), np.linalg.eigvalsh(res.Y[p])
        margin = min(ex[ex > 1e-5 * ex[-1]].min(), ey[ey > 1e-5 * ey[-1]].min())
        print(f"problem {p}: mu_ck {mu[p]:.1e}, margin {margin:.1e}: limit dC {eC:.2e} db {eb:.2e} | plain dC "
              f"{pC:.2e}", flush=True)
        errs.append(max(eC, eb))
        plain.append(pC)
    errs, plain = np.array(errs), np.array(plain)
    # the limit beats the plain adjoint on every member; near-degenerate members (small margin) converge
    # more slowly in mu, so the bar is on the median and the bulk
    assert len(errs) >= 12 and np.all(errs < plain)
    assert np.median(errs) <= 1e-5 and np.mean(errs <= 3e-5) >= 0.75
