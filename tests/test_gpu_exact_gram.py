"""Bit-exact Schur assembly on the reference's integer-pattern dense problem.

docs/evidence/bench/threading/benchmark_schur_scheduler.jl:23-54 builds
A_i[r, c] = (mod(17 i + 31 l + 13 r + 7 c, 41) - 20) / 23 (1-based, r <= c,
mirrored).  Scaled by 23 the entries are integers in [-20, 20], and at
X = Y = I (the cold start, interior_point.jl:710-721) the Schur matrix is the
Gram matrix <A_i, A_j>: every product and every partial sum is an integer
below 2^53, so any correct FP64 evaluation order gives it exactly.  The GPU
assembly (panel products with L_X^-1 = M_Y = I, split-K Gram, tile order,
ragged and triangular tiles, the flat strips of a short last tile row) must
therefore equal the CPU Gram bit for bit,
at the benchmark's own size too: a wrong index, tile offset or dropped partial
changes integers.  The FP32 and BF16x3 Gram kernels are exact on the same
data while every partial sum stays below 2^24 (k <= 130 here): integers up to
256 are exact in bf16, so the lo plane is zero and H'H carries the whole sum.
"""

import numpy as np
import pytest

import sdpx_oracle as O
from sdpx_mi355x import MI355XDenseBackend, ingest

from _helpers import rel_fro_lower


def scheduler_panel(k: int, m: int, block: int = 1) -> np.ndarray:
    """23 x the A_i of benchmark_schur_scheduler.jl:23-54, as an (m, k, k) array of integers."""
    i = np.arange(1, m + 1)[:, None, None]
    r = np.arange(1, k + 1)[None, :, None]
    c = np.arange(1, k + 1)[None, None, :]
    lo, hi = np.minimum(r, c), np.maximum(r, c)  # the formula is evaluated on the upper triangle, then mirrored
    return (np.mod(17 * i + 31 * block + 13 * lo + 7 * hi, 41) - 20).astype(np.float64)


def test_oracle_cold_start_gram_is_exact():
    k, m = 7, 12
    A = scheduler_panel(k, m)
    prob = O.ingest(np.ones(m), [A], [np.zeros((k, k))])
    S = O.schur_build(prob, [np.eye(k)], [np.eye(k)])
    V = A.reshape(m, k * k)
    assert np.array_equal(np.tril(S), np.tril(V @ V.T))


@pytest.mark.gpu
@pytest.mark.parametrize("k,m", [(65, 300), (130, 700), (400, 4000), (65, 160), (65, 276), (33, 20), (33, 64)])
def test_gpu_cold_start_gram_is_bit_exact(gpu_lib, k, m):
    """m mod 128 picks the shape of the last tile row: 32 x 256 strips (m = 160,
    276, 4000, 20: the last tile column even or odd), 64 x 128 strips (300, 700,
    64) or full tiles."""
    A = scheduler_panel(k, m)
    prob = ingest(np.ones(m), [A], [np.zeros((k, k))])
    b = MI355XDenseBackend(prob)
    try:
        I = np.eye(k)
        b.set_iterate(np.zeros(m), [I], [I])
        _, _, ok = b.compute_residuals(np.array([0.0]), affine=True, factor=True)
        assert ok
        b.schur_build()
        S = np.tril(b.get_schur())
    finally:
        b.close()
    V = A.reshape(m, k * k)
    ref = np.tril(V @ V.T)
    bad = np.argwhere(S != ref)
    print(f"k={k} m={m}: {len(bad)} entries differ; rel {rel_fro_lower(S, ref):.1e}", flush=True)
    assert len(bad) == 0, bad[:5]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [1, 2])
@pytest.mark.parametrize("k,m", [(65, 300), (130, 700)])
def test_gpu_reduced_precision_cold_start_gram_is_bit_exact(gpu_lib, k, m, precision):
    A = scheduler_panel(k, m)
    assert 400.0 * k * k < 2.0 ** 24  # every FP32 partial sum is an exact integer
    prob = ingest(np.ones(m), [A], [np.zeros((k, k))])
    b = MI355XDenseBackend(prob)
    try:
        I = np.eye(k)
        b.set_iterate(np.zeros(m), [I], [I])
        _, _, ok = b.compute_residuals(np.array([0.0]), affine=True, factor=True)
        assert ok
        b.set_precision(precision)
        b.schur_build()
        S = np.tril(b.get_schur())
    finally:
        b.close()
    V = A.reshape(m, k * k)
    ref = np.tril(V @ V.T)
    bad = np.argwhere(S != ref)
    print(f"precision {precision} k={k} m={m}: {len(bad)} entries differ", flush=True)
    assert len(bad) == 0, bad[:5]


@pytest.mark.gpu
@pytest.mark.parametrize("k,m", [(64, 300), (128, 700)])
def test_gpu_bf16x3_fp32_products_cold_start_gram_is_bit_exact(gpu_lib, k, m):
    """The BF16x3 assembly from the FP32-MFMA P-hat products (block sides divisible by 4): with
    X = Y = I the products reproduce the integer panel exactly and the Gram stays exact."""
    A = scheduler_panel(k, m)
    assert 400.0 * k * k < 2.0 ** 24  # every FP32 partial sum is an exact integer
    prob = ingest(np.ones(m), [A], [np.zeros((k, k))])
    b = MI355XDenseBackend(prob)
    try:
        I = np.eye(k)
        b.set_iterate(np.zeros(m), [I], [I])
        _, _, ok = b.compute_residuals(np.array([0.0]), affine=True, factor=True)
        assert ok
        b.set_precision(2)
        b.set_panel_precision(1)
        b.schur_build()
        S = np.tril(b.get_schur())
    finally:
        b.close()
    V = A.reshape(m, k * k)
    ref = np.tril(V @ V.T)
    bad = np.argwhere(S != ref)
    print(f"BF16x3 from FP32 products k={k} m={m}: {len(bad)} entries differ", flush=True)
    assert len(bad) == 0, bad[:5]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("kind", ["dense", "coo"])
def test_gpu_batched_cold_start_gram_is_bit_exact(gpu_lib, kind, precision):
    """Every padded block size of the batched panel kernels (k = 3 .. 64) and both
    input forms; members use the same integer pattern with their own block index.
    Precision 1 runs the FP32 panel output and the grouped FP32 Gram product: the
    entries are integers of at most 20 in magnitude, so every FP32 partial sum
    (at most 64^2 * 400 < 2^24) is exact and the FP32 route must give the same
    integers."""
    from sdpx_mi355x.batch import BatchSolver

    ks = [3, 9, 16, 17, 30, 32, 33, 47, 48, 49, 63, 64]
    ms = [5, 40, 60, 70, 130, 200, 150, 300, 260, 320, 400, 520]
    panels = [scheduler_panel(k, m, block=p + 1) for p, (k, m) in enumerate(zip(ks, ms))]
    with BatchSolver(ks, ms) as bs:
        if kind == "dense":
            # rows = vec(A_i), column-major (A_i is symmetric)
            bs.set_dense([A.reshape(A.shape[0], -1) for A in panels], [np.zeros((k, k)) for k in ks],
                         [np.ones(m) for m in ms])
        else:
            rows, cols, vals = [], [], []
            for A in panels:
                k = A.shape[1]
                iu = np.triu_indices(k)
                for Ai in A:
                    rows.append(iu[0].astype(np.int32))
                    cols.append(iu[1].astype(np.int32))
                    vals.append(Ai[iu])
            bs.set_coo(rows, cols, vals, [np.zeros((k, k)) for k in ks], [np.ones(m) for m in ms])
        bs.set_start([np.zeros(m) for m in ms], [np.eye(k) for k in ks], [np.eye(k) for k in ks])
        bs.set_precision(precision)
        assert bs.schur_at_start() == 0
        for p, A in enumerate(panels):
            V = A.reshape(A.shape[0], -1)
            ref = np.tril(V @ V.T)
            S = np.tril(bs.schur(p))
            bad = np.argwhere(S != ref)
            print(f"{kind} precision {precision} member {p}: k={ks[p]} m={ms[p]}: {len(bad)} entries differ",
                  flush=True)
            assert len(bad) == 0, (p, bad[:5])
