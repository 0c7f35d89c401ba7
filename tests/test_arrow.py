"""Block-arrow elimination of the Schur system (ArrowWorkspace, workspace.jl:318-400;
factor_arrow_kkt!, kkt.jl:2132-2349; solve_arrow_kkt!, kkt.jl:2977-3096): oracle pins on the CPU,
GPU parity through the C-ABI."""

import numpy as np
import pytest

import sdpx_oracle as O
from _helpers import random_pd
from sdpx_mi355x import backend as BK
from sdpx_mi355x._lib import SDPXError


def three_block_case():
    """test/sparse.jl:8-26 (sparse_block_data) with the X, Y of test/sparse.jl:146-155: m = 8, k = 2,
    active sets {1,2,5}, {1,3,6}, {1,4,7,8} (variable 1 global)."""
    m, k, L = 8, 2, 3
    A = [np.zeros((m, k, k)) for _ in range(L)]
    active = [[1, 2, 5], [1, 3, 6], [1, 4, 7, 8]]
    for l in range(1, L + 1):
        for p, i in enumerate(active[l - 1], 1):
            A[l - 1][i - 1, 0, 0] = 1 + p
            A[l - 1][i - 1, 1, 1] = 2 + l + p
            if (i + l) % 2 == 1:
                A[l - 1][i - 1, 0, 1] = A[l - 1][i - 1, 1, 0] = 0.1 * (i + l)
    X = [np.array([[2.0, 0.2], [0.2, 1.5]]), np.array([[1.8, -0.1], [-0.1, 2.2]]),
         np.array([[2.5, 0.3], [0.3, 1.7]])]
    Y = [np.array([[1.4, 0.1], [0.1, 1.9]]), np.array([[2.1, 0.2], [0.2, 1.3]]),
         np.array([[1.6, -0.2], [-0.2, 2.4]])]
    return np.ones(m), A, [np.zeros((k, k))] * L, X, Y


def arrow_model(blocks, shared, seed=11, k=2, per=1):
    """The structure of test/sparse.jl:444-458 (numpy's generator in place of Julia's): variables
    1..shared touch every block, then ``per`` local variables per block; random symmetric k x k
    coefficients, random symmetric C."""
    rng = np.random.default_rng(seed)
    m = shared + blocks * per
    A = [np.zeros((m, k, k)) for _ in range(blocks)]
    for l in range(blocks):
        for i in list(range(shared)) + [shared + l * per + j for j in range(per)]:
            T = rng.standard_normal((k, k))
            A[l][i] = (T + T.T) / 2
    C = []
    for _ in range(blocks):
        T = rng.standard_normal((k, k))
        C.append((T + T.T) / 2)
    return rng.standard_normal(m), A, C


def arrow_iterate(blocks, k=2, seed=7):
    rng = np.random.default_rng(seed)
    base_x = np.eye(k) * 2.0 + 0.2 * (np.ones((k, k)) - np.eye(k))
    base_y = np.eye(k) * 1.8 + 0.1 * (np.ones((k, k)) - np.eye(k))
    X, Y = [], []
    for _ in range(blocks):
        T = 0.1 * rng.standard_normal((k, k))
        X.append(base_x + (T + T.T) / 2)
        T = 0.1 * rng.standard_normal((k, k))
        Y.append(base_y + (T + T.T) / 2)
    return X, Y


def feasible_arrow_problem(blocks, shared, per, k, seed):
    """A strictly feasible problem of the arrow shape: X0 = Y0 = I at x0."""
    rng = np.random.default_rng(seed)
    m = shared + blocks * per
    A = [np.zeros((m, k, k)) for _ in range(blocks)]
    for l in range(blocks):
        for i in list(range(shared)) + [shared + l * per + j for j in range(per)]:
            T = rng.standard_normal((k, k)) / np.sqrt(k)
            A[l][i] = (T + T.T) / 2
    x0 = rng.standard_normal(m)
    C = [np.tensordot(x0, A[l], axes=1) - np.eye(k) for l in range(blocks)]
    c = np.array([sum(np.trace(A[l][i]) for l in range(blocks)) for i in range(m)])
    return c, A, C, x0


def _schur(prob, X, Y):
    return O.schur_build(prob, [O.kchol(b)[0] for b in X], [O.kchol(b)[0] for b in Y], Y)


# ---- oracle pins (CPU) --------------------------------------------------------------
def test_oracle_three_block_case():
    """test/sparse.jl:141-188: the arrow solve equals the dense solve (rtol 1e-11); variable 1 is the
    only global one."""
    c, A, C, X, Y = three_block_case()
    prob = O.ingest(c, A, C)
    gids, lids = O.arrow_partition(prob)
    assert list(gids) == [0] and [list(x) for x in lids] == [[1, 4], [2, 5], [3, 6, 7]]
    S = _schur(prob, X, Y)
    f, ok, att = O.factor_arrow_kkt(S, (gids, lids))
    assert ok and att == 0
    r = np.linspace(0.2, 1.6, 8)
    assert np.allclose(O.solve_arrow_kkt(f, r), np.linalg.solve(S, r), rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("blocks,shared", [(20, 4), (100, 12)])
def test_oracle_arrow_matches_dense_at_scale(blocks, shared):
    """test/sparse.jl:437-495 on the same structure: arrow dx = dense dx and S dx = rhs (the reference's
    bars are 1e-8; measured far below)."""
    c, A, C = arrow_model(blocks, shared)
    prob = O.ingest(c, A, C)
    part = O.arrow_partition(prob)
    assert part is not None and len(part[0]) == shared
    X, Y = arrow_iterate(blocks)
    S = _schur(prob, X, Y)
    f, ok, _ = O.factor_arrow_kkt(S, part)
    assert ok
    rhs = np.linspace(0.2, 1.6, prob.m)
    dxs = O.solve_arrow_kkt(f, rhs)
    dxd = O.solve_dense_kkt(O.factor_dense_kkt(S)[0], rhs)
    assert np.allclose(dxs, dxd, rtol=1e-10, atol=1e-12)
    assert np.allclose(S @ dxs, rhs, rtol=1e-10, atol=1e-12)


def test_oracle_local_ladder_closed_form():
    """A local block diag(1, -1e-7): the shift sqrt(eps) 10^(a-1) max(|D_ii|, 1) first clears
    1e-7 at a = 2, so the block takes 2 attempts; Sred (PD) none."""
    S = np.diag([1.0, -1e-7, 3.0, 2.0, 5.0])
    S[4, :4] = S[:4, 4] = [0.1, 0.0, 0.2, 0.3]
    gids, lids = np.array([4]), [np.array([0, 1]), np.array([2, 3])]
    f, ok, att = O.factor_arrow_kkt(S, (gids, lids))
    assert ok and att == 2
    f, ok, att = O.factor_arrow_kkt(np.diag([1.0, -1.0, 3.0, 2.0, 5.0]), (gids, lids))
    assert not ok and att == 6


def test_oracle_partition_rules():
    """workspace.jl:340-346: a variable on no block or all variables global gives no arrow route."""
    c, A, C, _, _ = three_block_case()
    A2 = [a.copy() for a in A]
    for a in A2:
        a[7] = 0.0
    assert O.arrow_partition(O.ingest(c, A2, C)) is None
    A3 = [a.copy() for a in A]
    for a in A3:
        a[:, 0, 0] += 1.0  # every variable on every block
    assert O.arrow_partition(O.ingest(c, A3, C)) is None


def test_oracle_arrow_route_solve_matches_dense_route():
    c, A, C, x0 = feasible_arrow_problem(12, 3, 2, 3, seed=5)
    prob = O.ingest(c, A, C)
    X0 = [np.eye(3)] * 12
    tol = 1e-8
    o = O.Options(eps_gap=tol, eps_primal=tol, eps_dual=tol)
    rd = O.solve(prob, o, x0=x0, X0=X0, Y0=X0)
    ra = O.solve(prob, O.Options(eps_gap=tol, eps_primal=tol, eps_dual=tol, kkt_route="block_arrow"), x0=x0,
                 X0=X0, Y0=X0)
    assert rd.status == ra.status == "Optimal"
    assert abs(ra.pobj - rd.pobj) <= 1e-8 * max(1.0, abs(rd.pobj))
    assert abs(ra.iterations - rd.iterations) <= 2


def test_host_owner_matches_oracle_partition():
    """The host's arrow_owner (the C-ABI's owner array) is the oracle's partition."""
    for case in (three_block_case()[:3], arrow_model(20, 4)):
        c, A, C = case
        part = O.arrow_partition(O.ingest(c, A, C))
        owner = BK.arrow_owner(BK.ingest(c, A, C))
        assert list(np.flatnonzero(owner < 0)) == list(part[0])
        for l, ids in enumerate(part[1]):
            assert list(np.flatnonzero(owner == l)) == list(ids)
    c, A, C = arrow_model(6, 2)
    sp = BK.ingest(c, A, C, sparse={0, 3})
    assert np.array_equal(BK.arrow_owner(sp), BK.arrow_owner(BK.ingest(c, A, C)))


# ---- GPU parity ---------------------------------------------------------------------
def _gpu_factor_solve(prob, x, X, Y, r, S=None):
    b = BK.MI355XDenseBackend(prob)
    try:
        b.set_arrow()
        b.set_iterate(x, X, Y)
        _, _, ok = b.compute_residuals([0.1] * prob.L, affine=True, factor=True)
        assert ok
        if S is None:
            b.schur_build()
        else:
            b.set_schur(S)
        fok, att = b.factorize()
        b.set_rhs(r)
        b.solve_direction()
        dx, _, _ = b.get_direction()
        return fok, att, dx, b.arrow_info()
    finally:
        b.close()


@pytest.mark.gpu
def test_gpu_three_block_case(gpu_lib):
    c, A, C, X, Y = three_block_case()
    r = np.linspace(0.2, 1.6, 8)
    ok, att, dx, info = _gpu_factor_solve(BK.ingest(c, A, C), np.zeros(8), X, Y, r)
    assert ok and att == 0 and info == (1, 3, 3)
    S = _schur(O.ingest(c, A, C), X, Y)
    assert np.allclose(dx, np.linalg.solve(S, r), rtol=1e-11, atol=1e-11)


@pytest.mark.gpu
@pytest.mark.parametrize("blocks,shared,per,k", [(20, 4, 1, 2), (100, 12, 1, 2), (40, 9, 5, 4),
                                                 (3, 20, 100, 16), (2, 7, 128, 16), (4, 0, 70, 12)])
def test_gpu_arrow_matches_oracle(gpu_lib, blocks, shared, per, k):
    """Local blocks of 1 to 128 variables (the largest in LDS above 64 KiB), with and without global
    variables: the device's factor and solve against the oracle's arrow elimination of the same S."""
    c, A, C = arrow_model(blocks, shared, seed=blocks + per, k=k, per=per)
    X, Y = arrow_iterate(blocks, k=k, seed=3)
    prob = O.ingest(c, A, C)
    S = _schur(prob, X, Y)
    r = np.linspace(0.2, 1.6, prob.m)
    ok, att, dx, info = _gpu_factor_solve(BK.ingest(c, A, C), np.zeros(prob.m), X, Y, r, S=S)
    f, oko, atto = O.factor_arrow_kkt(S, O.arrow_partition(prob))
    dxo = O.solve_arrow_kkt(f, r)
    err = np.max(np.abs(dx - dxo)) / np.max(np.abs(dxo))
    print(f"blocks={blocks} shared={shared} per={per} k={k}: attempts {att}/{atto} rel {err:.3e}", flush=True)
    assert ok == oko and att == atto and info[0] == shared and info[2] == per
    assert err <= 1e-10


@pytest.mark.gpu
def test_gpu_local_ladder(gpu_lib):
    """The planted local block diag(1, -1e-7) takes 2 attempts on the device as in the oracle, and the
    solve uses the shifted factor."""
    c, A, C, X, Y = three_block_case()
    prob = O.ingest(c, A, C)
    S = _schur(prob, X, Y)
    part = O.arrow_partition(prob)
    i, j = part[1][0]
    S[i, :] = S[:, i] = 0.0
    S[j, :] = S[:, j] = 0.0
    S[i, i], S[j, j] = 1.0, -1e-7
    r = np.linspace(0.2, 1.6, 8)
    ok, att, dx, _ = _gpu_factor_solve(BK.ingest(c, A, C), np.zeros(8), X, Y, r, S=S)
    f, oko, atto = O.factor_arrow_kkt(S, part)
    assert ok and oko and att == atto == 2
    dxo = O.solve_arrow_kkt(f, r)
    assert np.max(np.abs(dx - dxo)) <= 1e-9 * np.max(np.abs(dxo))


@pytest.mark.gpu
def test_gpu_arrow_solve_matches_oracle(gpu_lib):
    """End to end with kkt_route="block_arrow": status, objective 1e-8, iterations +-2 against the
    oracle on the same route, and against the device's dense route."""
    from sdpx_mi355x import SolverOptions, solve

    c, A, C, x0 = feasible_arrow_problem(30, 4, 3, 4, seed=8)
    X0 = [np.eye(4)] * 30
    tol = 1e-8
    ra = solve(BK.ingest(c, A, C), SolverOptions(eps_gap=tol, eps_primal=tol, eps_dual=tol, kkt_route="block_arrow"),
               x0=x0, X0=X0, Y0=X0)
    rd = solve(BK.ingest(c, A, C), SolverOptions(eps_gap=tol, eps_primal=tol, eps_dual=tol), x0=x0, X0=X0, Y0=X0)
    ref = O.solve(O.ingest(c, A, C), O.Options(eps_gap=tol, eps_primal=tol, eps_dual=tol, kkt_route="block_arrow"),
                  x0=x0, X0=X0, Y0=X0)
    print(f"arrow {ra.status} {ra.iterations} {ra.pobj:.12e} | dense {rd.status} {rd.iterations} | oracle "
          f"{ref.status} {ref.iterations} {ref.pobj:.12e}", flush=True)
    assert ref.status == "Optimal" and ra.status == ref.status == rd.status
    assert abs(ra.pobj - ref.pobj) <= 1e-8 * max(1.0, abs(ref.pobj))
    assert abs(ra.iterations - ref.iterations) <= 2


@pytest.mark.gpu
def test_gpu_arrow_guards(gpu_lib):
    c, A, C = arrow_model(5, 2)
    b = BK.MI355XDenseBackend(BK.ingest(c, A, C))
    try:
        assert b.arrow_info()[0] == -1
        b.set_arrow()
        with pytest.raises(SDPXError, match="FP64 only"):
            b.set_precision(1)
        with pytest.raises(SDPXError, match="no equality border"):
            b.set_equalities(np.ones((b.m, 1)))
        with pytest.raises(SDPXError, match="once per handle"):
            b.set_arrow()
    finally:
        b.close()
    c, A, C = arrow_model(2, 1, k=16, per=130)
    b = BK.MI355XDenseBackend(BK.ingest(c, A, C))
    try:
        with pytest.raises(SDPXError, match="at most 128"):
            b.set_arrow(np.where(np.arange(b.m) < 1, -1, (np.arange(b.m) - 1) // 130).astype(np.int32))
    finally:
        b.close()
