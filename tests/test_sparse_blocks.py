"""Sparse-block Schur route (SparseBlockCOO, sparse_schur_block!, schur.jl:566-640, 936-1032,
1229-1318): oracle pins on the CPU, GPU parity through the C-ABI."""

import math

import numpy as np
import pytest

import sdpx_oracle as O
from _helpers import golden, random_pd, rel_fro_lower
from sdpx_mi355x import backend as BK
from sdpx_mi355x._lib import SDPXError


def _mod1(x, d):
    return (x - 1) % d + 1


def outer_product_case():
    """test/sparse.jl:191-258: m = 7, k = 4, two blocks; X = 1.5 I + 0.03 J, Y = 1.2 I + 0.02 J."""
    m, d, nb = 7, 4, 2
    A = [np.zeros((m, d, d)) for _ in range(nb)]
    for block in range(1, nb + 1):
        for v in range(1, m + 1):
            r, c = _mod1(v + block, d) - 1, _mod1(2 * v + block, d) - 1
            A[block - 1][v - 1, r, c] += 0.2 * v
            A[block - 1][v - 1, c, r] += 0.2 * v
            A[block - 1][v - 1, r, r] += 1 + 0.1 * block
            if v % 2 == 0:
                A[block - 1][v - 1, c, c] += 0.3
    X = [1.5 * np.eye(d) + 0.03 * np.ones((d, d)) for _ in range(nb)]
    Y = [1.2 * np.eye(d) + 0.02 * np.ones((d, d)) for _ in range(nb)]
    return np.ones(m), A, [np.zeros((d, d))] * nb, X, Y


def sparse_mixed_case(seed, m, ks, sparse, density=0.15, touch=0.4):
    """Blocks listed in ``sparse`` get coefficients on a random subset of the variables (each
    A_i^(l) a few symmetric entries); the others are dense.  Feasible at x0 with X0 = I."""
    rng = np.random.default_rng(seed)
    A, Cs = [], []
    x0 = rng.standard_normal(m)
    for l, k in enumerate(ks):
        if l in sparse:
            Al = np.zeros((m, k, k))
            for i in np.flatnonzero(rng.random(m) < touch):
                mask = rng.random((k, k)) < density
                mask[rng.integers(k), rng.integers(k)] = True
                T = np.triu(np.where(mask, rng.standard_normal((k, k)), 0.0))
                Al[i] = T + np.triu(T, 1).T
        else:
            Al = rng.standard_normal((m, k, k))
            Al = (Al + Al.transpose(0, 2, 1)) / 2
        A.append(Al)
        Cs.append(np.tensordot(x0, Al, axes=1) - np.eye(k))
    Y0 = [np.eye(k) for k in ks]
    c = np.array([sum(np.trace(A[l][i] @ Y0[l]) for l in range(len(ks))) for i in range(m)])
    return c, A, Cs, x0


def _factors(X, Y):
    return [O.kchol(b)[0] for b in X], [O.kchol(b)[0] for b in Y]


# ---- oracle pins (CPU) --------------------------------------------------------------
def test_oracle_outer_product_case_equals_dense():
    """The reference's own case: sparse Schur == dense Schur, rtol 1e-12 (test/sparse.jl:251-257),
    schur_order ascending and equal to the active set (:206-216)."""
    c, A, C, X, Y = outer_product_case()
    dense, sp = O.ingest(c, A, C), O.ingest_sparse(c, A, C, blocks=[0, 1])
    for l in range(2):
        ids, _ = sp.sparse[l]
        assert np.all(np.diff(ids) > 0) and list(ids) == list(range(7))
    LX, MY = _factors(X, Y)
    Sd, Ss = O.schur_build(dense, LX, MY), O.schur_build(sp, LX, MY, Y)
    assert np.allclose(Ss, Sd, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("n", range(4))
def test_oracle_sparse_route_brute_force(n):
    """The sparse pair route reproduces the 50-digit brute force tr(A_i X^-1 A_j Y) of the tiny fixtures."""
    g = golden()
    A, X, Y, S = g[f"tiny{n}_A"], g[f"tiny{n}_X"], g[f"tiny{n}_Y"], g[f"tiny{n}_S"]
    m = A.shape[0]
    sp = O.ingest_sparse(np.zeros(m), [A], [np.zeros_like(X)], blocks=[0])
    Ss = O.schur_build(sp, [O.kchol(X)[0]], [O.kchol(Y)[0]], [Y])
    assert rel_fro_lower(Ss, S) <= 1e-13


def test_oracle_two_block_fixture_mixed_routes():
    """kkt_regressions.jl:488-565 data with block 2 sparse, block 1 dense: the 50-digit S."""
    g = golden()
    A = [g["kkt2_A1"], g["kkt2_A2"]]
    X, Y = [g["kkt2_X1"], g["kkt2_X2"]], [g["kkt2_Y1"], g["kkt2_Y2"]]
    sp = O.ingest_sparse(np.zeros(4), A, [np.zeros((2, 2))] * 2, blocks=[1])
    LX, MY = _factors(X, Y)
    assert rel_fro_lower(O.schur_build(sp, LX, MY, Y), g["kkt2_S"]) <= 1e-14


def test_oracle_planted_x_y_swap_is_caught():
    """W1 = X A_p Y^-1 (the factors' roles swapped) must fail the brute-force pin."""
    g = golden()
    A, X, Y, S = g["tiny2_A"], g["tiny2_X"], g["tiny2_Y"], g["tiny2_S"]
    sp = O.ingest_sparse(np.zeros(A.shape[0]), [A], [np.zeros_like(X)], blocks=[0])
    ids, coo = sp.sparse[0]
    bad = np.zeros_like(S)
    O.scatter_sparse_schur(bad, ids, O.sparse_schur_block(ids, coo, O.kchol(Y)[0], X))
    assert rel_fro_lower(bad, S) > 1e-3


def test_oracle_mixed_random_equals_dense():
    c, A, C, _ = sparse_mixed_case(5, 60, [6, 9, 4], sparse={0, 2})
    rng = np.random.default_rng(1)
    X = [random_pd(rng, k) for k in (6, 9, 4)]
    Y = [random_pd(rng, k) for k in (6, 9, 4)]
    LX, MY = _factors(X, Y)
    Sd = O.schur_build(O.ingest(c, A, C), LX, MY)
    Ss = O.schur_build(O.ingest_sparse(c, A, C, blocks=[0, 2]), LX, MY, Y)
    assert rel_fro_lower(Ss, Sd) <= 1e-13


def test_oracle_analytic_sparse_solve():
    """test/sparse.jl:330-352: the 2x2 analytic problem on the sparse route, p* = 2 sqrt 6."""
    A = np.zeros((2, 2, 2))
    A[0, 0, 0] = A[1, 1, 1] = 1.0
    Cm = np.array([[0.0, 1.0], [1.0, 0.0]])
    r = O.solve(O.ingest_sparse([2.0, 3.0], [A], [Cm], blocks=[0]), O.Options())
    assert r.status == "Optimal"
    assert abs(r.pobj - 2 * math.sqrt(6)) < 1e-8 and abs(r.dobj - 2 * math.sqrt(6)) < 1e-8


def test_host_packing_matches_oracle_coo():
    """The host's SparseBlockCOO.build / block_active (vectorised) give the arrays of the oracle's
    entry-by-entry build_block_coo (types/constraints.jl:79-109)."""
    c, A, C, _ = sparse_mixed_case(7, 40, [5, 8], sparse={0, 1})
    for Al in A:
        ids = BK.block_active(Al)
        assert list(ids) == list(O.block_active(Al))
        h, o = BK.SparseBlockCOO.build(Al, ids), O.build_block_coo(Al, ids)
        for f in ("ptr", "lin", "row", "col"):
            a, b = getattr(h, f), getattr(o, f)
            assert a.dtype == np.int32 and np.array_equal(a, b), f
        assert np.array_equal(h.val, o.val)
    prob = BK.ingest(c, A, C, sparse={1})
    assert prob.Av[1] is None and prob.Av[0] is not None and set(prob.sparse) == {1}


# ---- GPU parity ---------------------------------------------------------------------
def _gpu_schur(prob, x, X, Y):
    b = BK.MI355XDenseBackend(prob)
    try:
        b.set_iterate(x, X, Y)
        _, _, ok = b.compute_residuals([0.1] * prob.L, affine=True, factor=True)
        assert ok
        b.schur_build()
        return b.get_schur()
    finally:
        b.close()


@pytest.mark.gpu
def test_gpu_outer_product_case(gpu_lib):
    """test/sparse.jl:191-258 on the device: both blocks sparse (no dense block in S)."""
    c, A, C, X, Y = outer_product_case()
    S = _gpu_schur(BK.ingest(c, A, C, sparse={0, 1}), np.zeros(7), X, Y)
    LX, MY = _factors(X, Y)
    Sd = O.schur_build(O.ingest(c, A, C), LX, MY)
    assert np.allclose(np.tril(S), np.tril(Sd), rtol=1e-12, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("m,ks,sparse", [(60, [6, 9, 4], {0, 2}), (700, [12, 40, 70], {0, 1}),
                                         (300, [5, 33], {0, 1})])
def test_gpu_sparse_schur_matches_oracle(gpu_lib, m, ks, sparse):
    """Mixed dense + sparse blocks (and all-sparse), more than 256 active positions per block at
    m = 700: the device Schur against the oracle's sparse route, FP64 bar 1e-12."""
    c, A, C, x0 = sparse_mixed_case(m + len(ks), m, ks, sparse)
    rng = np.random.default_rng(m)
    X = [random_pd(rng, k) for k in ks]
    Y = [random_pd(rng, k) for k in ks]
    S = _gpu_schur(BK.ingest(c, A, C, sparse=sparse), x0, X, Y)
    LX, MY = _factors(X, Y)
    Sref = O.schur_build(O.ingest_sparse(c, A, C, blocks=sorted(sparse)), LX, MY, Y)
    print(f"m={m} ks={ks} sparse={sorted(sparse)} rel={rel_fro_lower(S, Sref):.3e}", flush=True)
    assert rel_fro_lower(S, Sref) <= 1e-12


@pytest.mark.gpu
def test_gpu_analytic_sparse_solve(gpu_lib):
    """test/sparse.jl:330-352 on the device: p* = d* = 2 sqrt 6."""
    from sdpx_mi355x import SolverOptions, solve

    A = np.zeros((2, 2, 2))
    A[0, 0, 0] = A[1, 1, 1] = 1.0
    Cm = np.array([[0.0, 1.0], [1.0, 0.0]])
    res = solve(BK.ingest([2.0, 3.0], [A], [Cm], sparse={0}), SolverOptions())
    assert res.status == "Optimal"
    assert abs(res.pobj - 2 * math.sqrt(6)) < 1e-8 and abs(res.dobj - 2 * math.sqrt(6)) < 1e-8


@pytest.mark.gpu
def test_gpu_sparse_solve_matches_oracle(gpu_lib):
    """A three-block problem with two sparse blocks solved end to end: status, objective 1e-8 and
    iterations +-2 against the oracle's sparse route."""
    from sdpx_mi355x import SolverOptions, solve

    ks, sparse = [6, 14, 5], {0, 2}
    c, A, C, x0 = sparse_mixed_case(11, 80, ks, sparse)
    tol = 1e-8
    X0 = [np.eye(k) for k in ks]
    res = solve(BK.ingest(c, A, C, sparse=sparse), SolverOptions(eps_gap=tol, eps_primal=tol, eps_dual=tol),
                x0=x0, X0=X0, Y0=X0)
    ref = O.solve(O.ingest_sparse(c, A, C, blocks=sorted(sparse)),
                  O.Options(eps_gap=tol, eps_primal=tol, eps_dual=tol), x0=x0, X0=X0, Y0=X0)
    print(f"gpu {res.status} {res.iterations} {res.pobj:.12e} | oracle {ref.status} {ref.iterations} "
          f"{ref.pobj:.12e}", flush=True)
    assert ref.status == "Optimal" and res.status == ref.status
    assert abs(res.pobj - ref.pobj) <= 1e-8 * max(1.0, abs(ref.pobj))
    assert abs(res.iterations - ref.iterations) <= 2


@pytest.mark.gpu
def test_gpu_sparse_block_guards(gpu_lib):
    """A sparse block assembles in FP64 as given: reduced precision, equilibration and a dense
    re-upload of that block are rejected with the library's error."""
    c, A, C, _ = sparse_mixed_case(3, 30, [4, 5], {1})
    b = BK.MI355XDenseBackend(BK.ingest(c, A, C, sparse={1}))
    try:
        with pytest.raises(SDPXError, match="FP64 only"):
            b.set_precision(1)
        with pytest.raises(SDPXError, match="sparse blocks"):
            b.equilibrate()
        with pytest.raises(SDPXError, match="sparse coordinates"):
            Av = np.zeros((30, 25))
            BK.check(b.lib.sdpx_dense_set_block(b.h, 1, BK.dptr(Av), BK.dptr(np.zeros(25))), "set_block")
    finally:
        b.close()
