"""Termination statuses that rounding decides, on the two full batched workloads.

The reference's adaptive controller compares residuals that sit at rounding
level on these feasible-start problems: its feasibility-growth test divides two
such residuals (adaptive_parameters.jl:550-586, fed the raw p_res and d_res at
interior_point.jl:2254-2276), and near the gap target the PSD-margin trial of
the full step (generic.jl:318-340) is taken between values that differ in the
last bits.  On a few members the status therefore depends on the summation
order of the A-contractions, which Float64 leaves open: OpenBLAS dgemv (the
reference's call) is one valid order, a pairwise or strided reduction another.

Each case below is a member of the 1024-problem layer batch (dense n=32, m=128,
seed 0, gap 1e-9) or of the 4096-problem uneven sparse batch (seed 0, gap 1e-8)
that the GPU's batched route ends non-Optimal.  The oracle solves it to Optimal
with the BLAS order and ends it with the GPU's status under the named order.
The GPU side of the check is tests/test_gpu_batch.py.
"""

import numpy as np
import pytest

import sdpx_oracle as O

from sdpx_mi355x import generators as G

# (workload, member, contract order, build order, status the oracle reaches with that order)
ROUNDING_DECIDED = [
    ("layer", 867, "strided:8", "blas", "NumericalBreakdown"),
    ("layer", 881, "strided:16", "blas", "Stalled"),
    ("layer", 890, "pairwise", "blas", "NumericalBreakdown"),
    ("sparse", 1727, "blas", "pairwise", "Stalled"),
]


def member(workload: str, p: int):
    """Member p of the full batched workload, its oracle problem and tolerance."""
    if workload == "layer":
        g = G.batched_dense_layer(p + 1, seed=0)[p]
        return g, O.ingest(g.c, [g.A_blocks()], [g.C]), 1e-9
    g = G.batched_sparse(4096, seed=0, only=[p])[0]
    return g, O.ingest(g.c, [g.dense_rows().reshape(g.m, g.k, g.k)], [g.C]), 1e-8


def oracle_solve(workload, p, contract="blas", build="blas"):
    g, prob, tol = member(workload, p)
    o = O.Options(eps_gap=tol, eps_primal=tol, eps_dual=tol)
    with O.summation_order(contract, build):
        return O.solve(prob, o, x0=g.x0, X0=[np.eye(g.k)], Y0=[np.eye(g.k)])


def test_layer_member_generator_matches_batch_prefix():
    full = G.batched_dense_layer(900, seed=0)
    one = G.batched_dense_layer(868, seed=0)[867]
    assert np.array_equal(full[867].Av_rows, one.Av_rows) and np.array_equal(full[867].C, one.C)
    assert np.array_equal(full[867].c, one.c)


def test_sparse_member_generator_matches_batch():
    full = G.batched_sparse(60, seed=0)
    for p in (0, 17, 59):
        g = G.batched_sparse(60, seed=0, only=[p])[0]
        assert (g.k, g.m) == (full[p].k, full[p].m)
        for i in range(g.m):
            assert np.array_equal(g.rows[i], full[p].rows[i]) and np.array_equal(g.vals[i], full[p].vals[i])


def test_summation_orders_agree_to_rounding():
    g, prob, _ = member("layer", 3)
    rng = np.random.default_rng(0)
    Y = rng.standard_normal((g.k, g.k))
    x = rng.standard_normal(g.m)
    ref_c, ref_b = O.contract(prob, 0, Y), O.build_P(prob, 0, x)
    for order in ("pairwise", "strided:8", "strided:16"):
        with O.summation_order(order, order):
            c, b = O.contract(prob, 0, Y), O.build_P(prob, 0, x)
        assert np.max(np.abs(c - ref_c)) <= 64 * O.EPS * np.max(np.abs(ref_c))
        assert np.max(np.abs(b - ref_b)) <= 64 * O.EPS * np.max(np.abs(ref_b))
    with pytest.raises(ValueError):
        O.summation_order("strided:0")


@pytest.mark.parametrize("workload,p,contract,build,status", ROUNDING_DECIDED)
def test_status_decided_by_summation_order(workload, p, contract, build, status):
    blas = oracle_solve(workload, p)
    other = oracle_solve(workload, p, contract, build)
    print(f"{workload}[{p}]: blas {blas.status}/{blas.iterations}, {contract}/{build} {other.status}/{other.iterations}")
    assert blas.status == "Optimal"
    assert other.status == status
