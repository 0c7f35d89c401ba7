"""Batched route: every member against the single-problem oracle, through the C-ABI.

There is no batched solver in the reference, so each member is pinned
individually: its Schur matrix against the oracle's schur_build at the same
point, its solve against the oracle's solve of that member alone (status,
objective 1e-8 relative, iterations +-2, certificate).
"""

import numpy as np
import pytest

import sdpx_oracle as O

from sdpx_mi355x import generators as G
from sdpx_mi355x.batch import BatchSolver
from sdpx_mi355x.policy import SolverOptions

from _helpers import rel_fro_lower


def _oracle_problem(g):
    A = g.A_blocks() if isinstance(g, G.DenseSDP) else g.dense_rows().reshape(g.m, g.k, g.k)
    return O.ingest(g.c, [A], [g.C])


def _load(bs, probs):
    dense = [g for g in probs if isinstance(g, G.DenseSDP)]
    if len(dense) == len(probs):
        bs.set_dense([g.Av_rows for g in probs], [g.C for g in probs], [g.c for g in probs])
    else:
        rows, cols, vals = [], [], []
        for g in probs:
            rows += list(g.rows)
            cols += list(g.cols)
            vals += list(g.vals)
        bs.set_coo(rows, cols, vals, [g.C for g in probs], [g.c for g in probs])


def _random_pd(rng, k):
    Q = rng.standard_normal((k, k))
    return Q @ Q.T / k + 0.5 * np.eye(k)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["dense", "coo"])
def test_batch_schur_matches_oracle(gpu_lib, kind):
    rng = np.random.default_rng(3)
    ks = [3, 8, 17, 33, 64, 1, 2, 40]
    ms = [7, 20, 50, 70, 130, 3, 5, 300]
    probs = [G.dense_feasible(k, m, seed=10 + i) if kind == "dense" else G.sparse_feasible(max(k, 1), m, seed=10 + i)
             for i, (k, m) in enumerate(zip(ks, ms))]
    Xs = [_random_pd(rng, k) for k in ks]
    Ys = [_random_pd(rng, k) for k in ks]
    with BatchSolver(ks, ms) as bs:
        _load(bs, probs)
        bs.set_start([g.x0 for g in probs], Xs, Ys)
        assert bs.schur_at_start() == 0
        for p, g in enumerate(probs):
            prob = _oracle_problem(g)
            LX, _ = O.kchol(Xs[p])
            MY, _ = O.kchol(Ys[p])
            ref = O.schur_build(prob, [LX], [MY])
            got = bs.schur(p)
            print(f"member {p}: k={ks[p]} m={ms[p]} rel={rel_fro_lower(got, ref):.2e}", flush=True)
            assert rel_fro_lower(got, ref) <= 1e-12


@pytest.mark.gpu
def test_batch_autotune_orders_give_identical_schur(gpu_lib):
    """The autotuner's candidate tile orders of the grouped Gram (K-descending, member-major, XCD
    bins) only change which work-group computes a tile: every member's Schur matrix is bitwise the
    same, and the chosen order is one of the timed ones."""
    rng = np.random.default_rng(8)
    ks = [3, 8, 17, 33, 64, 40, 12, 25, 5, 50]
    ms = [7, 20, 50, 70, 130, 300, 40, 90, 12, 200]
    probs = [G.dense_feasible(k, m, seed=40 + i) for i, (k, m) in enumerate(zip(ks, ms))]
    Xs = [_random_pd(rng, k) for k in ks]
    Ys = [_random_pd(rng, k) for k in ks]
    with BatchSolver(ks, ms) as bs:
        _load(bs, probs)
        bs.set_start([g.x0 for g in probs], Xs, Ys)
        tuned = bs.autotune(repeats=2)
        print(tuned, flush=True)
        assert tuned["gram_order"] in BatchSolver.GRAM_ORDERS and all(v > 0 for v in tuned["gram_ms"].values())
        ref = None
        for order in BatchSolver.GRAM_ORDERS:
            bs.set_gram_order(order)
            assert bs.schur_at_start() == 0
            S = [bs.schur(p) for p in range(len(ks))]
            if ref is None:
                ref = S
            assert all(np.array_equal(np.tril(a), np.tril(b)) for a, b in zip(S, ref)), order


def _check_members(res, probs, o, oo, max_iter_diff=2):
    for p, g in enumerate(probs):
        prob = _oracle_problem(g)
        X0 = np.eye(g.k)
        ref = O.solve(prob, oo, x0=g.x0, X0=[X0], Y0=[X0])
        print(f"member {p}: k={g.k} m={g.m} gpu {res.status[p]} it={res.iterations[p]} pobj={res.pobj[p]:.10e} | "
              f"oracle {ref.status} it={ref.iterations} pobj={ref.pobj:.10e}", flush=True)
        assert res.status[p] == ref.status
        if ref.status != "Optimal":
            continue
        assert abs(res.pobj[p] - ref.pobj) <= 1e-8 * max(1.0, abs(ref.pobj))
        assert abs(int(res.iterations[p]) - ref.iterations) <= max_iter_diff
        r = O.Result(res.status[p], res.x[p], [res.X[p]], [res.Y[p]], res.pobj[p], res.dobj[p],
                     int(res.iterations[p]), 0.0, 0.0, float(res.gap_rel[p]))
        cert = O.certificate(prob, r, o.eps_primal, o.eps_dual, o.eps_gap)
        assert cert["valid"], (p, cert)


@pytest.mark.gpu
@pytest.mark.parametrize("strategy", ["adaptive", "fixed"])
def test_batch_dense_members_match_oracle(gpu_lib, strategy):
    ks = [5, 12, 20, 32, 9, 3, 2, 1]
    ms = [12, 40, 60, 128, 30, 6, 3, 2]
    probs = [G.dense_feasible(k, m, seed=100 + i) for i, (k, m) in enumerate(zip(ks, ms))]
    tol = 1e-8
    o = SolverOptions(eps_gap=tol, eps_primal=tol, eps_dual=tol, parameter_strategy=strategy)
    oo = O.Options(eps_gap=tol, eps_primal=tol, eps_dual=tol, parameter_strategy=strategy)
    with BatchSolver(ks, ms) as bs:
        _load(bs, probs)
        bs.set_start([g.x0 for g in probs], [np.eye(k) for k in ks], [np.eye(k) for k in ks])
        res = bs.solve(o)
    _check_members(res, probs, o, oo)


@pytest.mark.gpu
def test_batch_sparse_uneven_members_match_oracle(gpu_lib):
    """A small draw of the uneven sparse workload (k in [8, 24] here), COO input densified on the device."""
    probs = G.batched_sparse(6, seed=5, kmin=8, kmax=24)
    ks = [g.k for g in probs]
    ms = [g.m for g in probs]
    tol = 1e-8
    o = SolverOptions(eps_gap=tol, eps_primal=tol, eps_dual=tol)
    oo = O.Options(eps_gap=tol, eps_primal=tol, eps_dual=tol)
    with BatchSolver(ks, ms) as bs:
        _load(bs, probs)
        bs.set_start([g.x0 for g in probs], [np.eye(k) for k in ks], [np.eye(k) for k in ks])
        res = bs.solve(o)
    _check_members(res, probs, o, oo)


@pytest.mark.gpu
def test_batch_layer_members_match_oracle(gpu_lib):
    """The differentiable-layer shape (k = 32, m = 128, perturbed C and b) at gap 1e-9."""
    probs = G.batched_dense_layer(4, seed=2)
    tol = 1e-9
    o = SolverOptions(eps_gap=tol, eps_primal=tol, eps_dual=tol)
    oo = O.Options(eps_gap=tol, eps_primal=tol, eps_dual=tol)
    with BatchSolver([32] * 4, [128] * 4) as bs:
        _load(bs, probs)
        bs.set_start([g.x0 for g in probs], [np.eye(32)] * 4, [np.eye(32)] * 4)
        res = bs.solve(o)
    _check_members(res, probs, o, oo)


@pytest.mark.gpu
def test_batch_member_failure_is_isolated(gpu_lib):
    """A member whose start is not positive definite breaks down alone; the others solve."""
    ks, ms = [6, 6, 10], [15, 15, 30]
    probs = [G.dense_feasible(k, m, seed=200 + i) for i, (k, m) in enumerate(zip(ks, ms))]
    X0 = [np.eye(k) for k in ks]
    X0[1] = -np.eye(6)
    o = SolverOptions(eps_gap=1e-8, eps_primal=1e-8, eps_dual=1e-8)
    with BatchSolver(ks, ms) as bs:
        _load(bs, probs)
        bs.set_start([g.x0 for g in probs], X0, [np.eye(k) for k in ks])
        res = bs.solve(o)
        reasons = bs.reasons()
    assert res.status[1] == "NumericalBreakdown"
    assert res.status[0] == res.status[2] == "Optimal"
    assert reasons == ["optimal", "block_factor_failed", "optimal"]


@pytest.mark.gpu
def test_batch_matches_single_problem_gpu_path(gpu_lib):
    """Same member through the batched route and the single-problem dense route."""
    from sdpx_mi355x import ingest, solve

    g = G.dense_feasible(24, 90, seed=7)
    o = SolverOptions(eps_gap=1e-8, eps_primal=1e-8, eps_dual=1e-8)
    single = solve(ingest(g.c, [g.A_blocks()], [g.C]), o, x0=g.x0, X0=[g.X0], Y0=[g.Y0])
    with BatchSolver([24, 24], [90, 90]) as bs:
        _load(bs, [g, g])
        bs.set_start([g.x0, g.x0], [g.X0, g.X0], [g.Y0, g.Y0])
        res = bs.solve(o)
    assert res.status == ["Optimal", "Optimal"]
    assert res.iterations[0] == res.iterations[1]
    assert res.pobj[0] == res.pobj[1]  # members are independent and deterministic
    # the two routes round differently (fused vs grouped kernels), and the adaptive
    # controller turns rounding-level differences into different step choices: the
    # oracle itself takes 10 or 11 iterations on this problem depending on the
    # summation order of its A-contractions (sdpx_oracle.summation_order), so the
    # tolerance is the one each route meets against the oracle (+-2)
    assert abs(int(res.iterations[0]) - single.iterations) <= 2
    assert abs(res.pobj[0] - single.pobj) <= 1e-8 * max(1.0, abs(single.pobj))
    # with the fixed strategy no decision sits at rounding level: the routes take the same steps
    of = SolverOptions(eps_gap=1e-8, eps_primal=1e-8, eps_dual=1e-8, parameter_strategy="fixed")
    single_f = solve(ingest(g.c, [g.A_blocks()], [g.C]), of, x0=g.x0, X0=[g.X0], Y0=[g.Y0])
    with BatchSolver([24], [90]) as bs:
        _load(bs, [g])
        bs.set_start([g.x0], [g.X0], [g.Y0])
        rf = bs.solve(of)
    assert rf.status[0] == single_f.status == "Optimal"
    assert int(rf.iterations[0]) == single_f.iterations
    assert abs(rf.pobj[0] - single_f.pobj) <= 1e-10 * max(1.0, abs(single_f.pobj))


@pytest.mark.gpu
def test_batch_coo_gather_matches_dense_panel(gpu_lib):
    """COO input runs Av x and Av' w as sparse gathers (duplicates summed in entry order);
    the same problems given dense run the dense panel passes.  With the fixed
    strategy both trajectories agree to rounding."""
    probs = G.batched_sparse(5, seed=9, kmin=6, kmax=20)
    ks = [g.k for g in probs]
    ms = [g.m for g in probs]
    rows, cols, vals = [], [], []
    for g in probs:
        for i in range(g.m):
            # split the first entry of every constraint into two exact halves (a duplicate position)
            r, c, v = list(g.rows[i]), list(g.cols[i]), list(g.vals[i])
            rows.append(np.array(r + [r[0]], dtype=np.int32))
            cols.append(np.array(c + [c[0]], dtype=np.int32))
            vals.append(np.array([v[0] / 2] + v[1:] + [v[0] / 2]))
    tol = 1e-8
    o = SolverOptions(eps_gap=tol, eps_primal=tol, eps_dual=tol, parameter_strategy="fixed")
    eyes = [np.eye(k) for k in ks]
    with BatchSolver(ks, ms) as bs:
        bs.set_coo(rows, cols, vals, [g.C for g in probs], [g.c for g in probs])
        bs.set_start([g.x0 for g in probs], eyes, eyes)
        res_s = bs.solve(o)
    with BatchSolver(ks, ms) as bs:
        bs.set_dense([g.dense_rows() for g in probs], [g.C for g in probs], [g.c for g in probs])
        bs.set_start([g.x0 for g in probs], eyes, eyes)
        res_d = bs.solve(o)
    for p in range(len(probs)):
        print(f"member {p}: coo {res_s.status[p]} it={res_s.iterations[p]} pobj={res_s.pobj[p]:.12e} | "
              f"dense {res_d.status[p]} it={res_d.iterations[p]} pobj={res_d.pobj[p]:.12e}", flush=True)
        assert res_s.status[p] == res_d.status[p] == "Optimal"
        assert abs(int(res_s.iterations[p]) - int(res_d.iterations[p])) <= 1
        assert abs(res_s.pobj[p] - res_d.pobj[p]) <= 1e-9 * max(1.0, abs(res_d.pobj[p]))


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["dense", "coo"])
def test_batch_fp32_schur_close_to_oracle(gpu_lib, kind):
    """Mixed-precision assembly: Phat rounded to FP32 by the panel pass and the grouped FP32 Gram
    product on FP32 MFMA.  The Schur matrix differs from the oracle's FP64 one by FP32 rounding only
    (relative Frobenius <= 2e-6 here; the same members assemble to 1e-12 in FP64), and it is not
    the FP64 matrix (the FP32 route really ran)."""
    rng = np.random.default_rng(5)
    ks = [3, 8, 17, 33, 64, 1, 2, 40]
    ms = [7, 20, 50, 70, 130, 3, 5, 300]
    probs = [G.dense_feasible(k, m, seed=10 + i) if kind == "dense" else G.sparse_feasible(max(k, 1), m, seed=10 + i)
             for i, (k, m) in enumerate(zip(ks, ms))]
    Xs = [_random_pd(rng, k) for k in ks]
    Ys = [_random_pd(rng, k) for k in ks]
    with BatchSolver(ks, ms) as bs:
        _load(bs, probs)
        bs.set_start([g.x0 for g in probs], Xs, Ys)
        bs.set_precision(1)
        assert bs.schur_at_start() == 0
        S32 = [bs.schur(p) for p in range(len(ks))]
        bs.set_precision(0)
        assert bs.schur_at_start() == 0
        S64 = [bs.schur(p) for p in range(len(ks))]
    for p, g in enumerate(probs):
        prob = _oracle_problem(g)
        LX, _ = O.kchol(Xs[p])
        MY, _ = O.kchol(Ys[p])
        ref = O.schur_build(prob, [LX], [MY])
        e32, e64 = rel_fro_lower(S32[p], ref), rel_fro_lower(S64[p], ref)
        print(f"member {p}: k={ks[p]} m={ms[p]} fp32 rel={e32:.2e} fp64 rel={e64:.2e}", flush=True)
        assert e32 <= 2e-6 and e64 <= 1e-12
        if ks[p] > 2:
            assert e32 > 1e-12  # rounded to FP32 (k <= 2 panels can be exact in FP32)


def _mixed_options(tol, **kw):
    from sdpx_mi355x.batch import BatchOptions

    return BatchOptions.from_solver_options(SolverOptions(eps_gap=tol, eps_primal=tol, eps_dual=tol,
                                                          precision="mixed", **kw))


@pytest.mark.gpu
@pytest.mark.parametrize("budget", [2, 32])
@pytest.mark.parametrize("workload", ["dense", "sparse", "layer"])
def test_batch_mixed_members_match_oracle(gpu_lib, workload, budget):
    """Mixed precision in the batched route: FP32 Schur assembly, every direction refined against the
    matrix-free FP64 operator.  Each member reaches the FP64 oracle's optimum (status, objective 1e-8
    relative, iterations +-2, certificate) with the batched default refinement budget (2) and the
    reference's (32); with 32 the reduced-precision Schur matrix carries most steps."""
    if workload == "dense":
        ks = [5, 12, 20, 32, 9, 3, 2, 1]
        ms = [12, 40, 60, 128, 30, 6, 3, 2]
        probs = [G.dense_feasible(k, m, seed=100 + i) for i, (k, m) in enumerate(zip(ks, ms))]
        tol = 1e-8
    elif workload == "sparse":
        probs = G.batched_sparse(6, seed=5, kmin=8, kmax=24)
        tol = 1e-8
    else:
        probs = G.batched_dense_layer(4, seed=2)
        tol = 1e-9
    ks = [g.k for g in probs]
    ms = [g.m for g in probs]
    o = SolverOptions(eps_gap=tol, eps_primal=tol, eps_dual=tol)
    oo = O.Options(eps_gap=tol, eps_primal=tol, eps_dual=tol)
    with BatchSolver(ks, ms) as bs:
        _load(bs, probs)
        bs.set_start([g.x0 for g in probs], [np.eye(k) for k in ks], [np.eye(k) for k in ks])
        om = _mixed_options(tol)
        om.mixed_max_refine = budget
        res = bs.solve(om)
        reduced = bs.mixed_iterations()
    print(f"reduced-precision steps per member: {reduced.tolist()} of {res.iterations.tolist()}", flush=True)
    _check_members(res, probs, o, oo)
    big = [p for p in range(len(ks)) if ks[p] >= 5]
    assert all(reduced[p] >= 1 for p in big)
    if budget == 32:
        assert sum(int(reduced[p]) for p in big) >= 0.5 * sum(int(res.iterations[p]) for p in big)


@pytest.mark.gpu
def test_batch_mixed_guard_fallback_is_the_fp64_solve(gpu_lib):
    """A pivot-ratio limit no Schur matrix meets sends every member back to FP64 at its first step,
    before any reduced-precision direction is used: the solve is then the FP64 solve bit for bit
    (the redone step recomputes the block factors and rebuilds S in FP64)."""
    probs = G.batched_sparse(5, seed=11, kmin=6, kmax=20)
    ks = [g.k for g in probs]
    ms = [g.m for g in probs]
    tol = 1e-8
    eyes = [np.eye(k) for k in ks]
    with BatchSolver(ks, ms) as bs:
        _load(bs, probs)
        bs.set_start([g.x0 for g in probs], eyes, eyes)
        r64 = bs.solve(SolverOptions(eps_gap=tol, eps_primal=tol, eps_dual=tol))
        om = _mixed_options(tol)
        om.cond_limit = 0.5  # below every (max L_ii / min L_ii)^2 >= 1
        bs.set_start([g.x0 for g in probs], eyes, eyes)
        rm = bs.solve(om)
        reduced = bs.mixed_iterations()
        # no refinement steps allowed: every member fails the predictor or corrector guard
        oz = _mixed_options(tol)
        oz.mixed_max_refine = 0
        bs.set_start([g.x0 for g in probs], eyes, eyes)
        rz = bs.solve(oz)
        reduced_z = bs.mixed_iterations()
    assert reduced.tolist() == [0] * len(ks)
    assert rm.status == r64.status
    assert rm.iterations.tolist() == r64.iterations.tolist()
    assert np.array_equal(rm.pobj, r64.pobj) and np.array_equal(rm.dobj, r64.dobj)
    assert reduced_z.tolist() == [0] * len(ks)
    assert all(s == "Optimal" for s in rz.status)
    assert np.allclose(rz.pobj, r64.pobj, rtol=1e-8, atol=1e-8)


def _full_batch(workload):
    if workload == "layer":
        probs, tol = G.batched_dense_layer(1024, seed=0), 1e-9
    else:
        probs, tol = G.batched_sparse(4096, seed=0), 1e-8
    return probs, SolverOptions(eps_gap=tol, eps_primal=tol, eps_dual=tol)


@pytest.mark.gpu
@pytest.mark.parametrize("workload", ["layer", "sparse"])
def test_full_batch_status_matches_oracle(gpu_lib, workload):
    """Every member of the full workload ends with the oracle's status.

    A member the GPU ends non-Optimal must end so in the oracle too: with the
    BLAS summation order, or with the order test_status_rounding.py names for
    it (there the oracle's own status changes with the summation order).  A
    sample of Optimal members is checked against the oracle's objective."""
    from test_status_rounding import ROUNDING_DECIDED, oracle_solve

    probs, o = _full_batch(workload)
    ks, ms = [g.k for g in probs], [g.m for g in probs]
    with BatchSolver(ks, ms) as bs:
        _load(bs, probs)
        bs.set_start([g.x0 for g in probs], [np.eye(k) for k in ks], [np.eye(k) for k in ks])
        res = bs.solve(o)
    bad = [p for p, s in enumerate(res.status) if s != "Optimal"]
    print(f"{workload}: non-Optimal members {[(p, res.status[p], int(res.iterations[p])) for p in bad]}", flush=True)
    named = {p: (c, b) for w, p, c, b, _ in ROUNDING_DECIDED if w == workload}
    for p in bad:
        ref = oracle_solve(workload, p)
        statuses = {ref.status}
        if p in named:
            statuses.add(oracle_solve(workload, p, *named[p]).status)
        print(f"member {p}: gpu {res.status[p]}, oracle statuses {sorted(statuses)}", flush=True)
        assert res.status[p] in statuses
    rng = np.random.default_rng(5)
    oo = O.Options(eps_gap=o.eps_gap, eps_primal=o.eps_primal, eps_dual=o.eps_dual)
    good = [p for p in range(len(probs)) if res.status[p] == "Optimal"]
    for p in rng.choice(good, size=8, replace=False):
        g = probs[p]
        ref = O.solve(_oracle_problem(g), oo, x0=g.x0, X0=[np.eye(g.k)], Y0=[np.eye(g.k)])
        print(f"member {p}: k={g.k} m={g.m} gpu it={res.iterations[p]} oracle {ref.status} it={ref.iterations}",
              flush=True)
        assert ref.status == "Optimal"
        assert abs(res.pobj[p] - ref.pobj) <= 1e-8 * max(1.0, abs(ref.pobj))
    assert np.all(res.gap_rel[good] <= o.eps_gap)


def _block_coo_of(g):
    """Member g's SparseBlockCOO as the reference stores it (build_block_coo, types/constraints.jl:79-109)."""
    D = g.dense_rows().reshape(g.m, g.k, g.k)
    ptr, lin, row, col, val = [1], [], [], [], []
    for i in range(g.m):
        for c in range(g.k):
            for r in np.nonzero(D[i][:, c])[0]:
                row.append(r + 1)
                col.append(c + 1)
                lin.append(c * g.k + r + 1)
                val.append(D[i][r, c])
        ptr.append(len(row) + 1)
    return ptr, lin, row, col, val


@pytest.mark.gpu
def test_batch_block_coo_as_stored_matches_coo(gpu_lib):
    """sdpx_batch_set_block_coo (SparseBlockCOO: 1-based, both triangles, lin) loads the same
    constraints as sdpx_batch_set_coo: the Schur matrices at the start are bitwise equal."""
    probs = G.batched_sparse(6, seed=4, kmin=5, kmax=24)
    ks, ms = [g.k for g in probs], [g.m for g in probs]
    rng = np.random.default_rng(2)
    Xs = [_random_pd(rng, k) for k in ks]
    Ys = [_random_pd(rng, k) for k in ks]
    S = {}
    for how in ("coo", "block_coo"):
        with BatchSolver(ks, ms) as bs:
            if how == "coo":
                _load(bs, probs)
            else:
                bs.set_block_coo([_block_coo_of(g) for g in probs], [g.C for g in probs], [g.c for g in probs])
            bs.set_start([g.x0 for g in probs], Xs, Ys)
            assert bs.schur_at_start() == 0
            S[how] = [bs.schur(p) for p in range(len(probs))]
    for p in range(len(probs)):
        assert np.array_equal(np.tril(S["coo"][p]), np.tril(S["block_coo"][p])), p
