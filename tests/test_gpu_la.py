"""The LA provider seam's host-buffer operations (la_backend.jl:1529-1769) on the device, through the
C-ABI, against LAPACK / BLAS (numpy, scipy) on the same operands: values to 1e-12, the reference's
in-place conventions (which triangle is written, what is kept), failure statuses."""

import numpy as np
import pytest
import scipy.linalg as sla

from sdpx_mi355x.la import MI355XLABackend

pytestmark = pytest.mark.gpu


def _spd(rng, n):
    Q = rng.standard_normal((n, n))
    return np.asfortranarray(Q @ Q.T / n + np.eye(n))


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def la(gpu_lib):
    with MI355XLABackend() as b:
        yield b


@pytest.mark.parametrize("n", [1, 63, 300])
def test_la_chol(la, n):
    """cholesky!(Symmetric(A, :L)): the lower triangle becomes L, the upper one is kept; a view with a
    leading dimension above n works."""
    rng = np.random.default_rng(n)
    S = _spd(rng, n)
    big = np.asfortranarray(rng.standard_normal((n + 5, n)))
    A = big[:n, :]
    A[np.tril_indices(n)] = S[np.tril_indices(n)]
    upper = np.triu(A, 1).copy()
    assert la.la_chol(A)
    assert _rel(np.tril(A), np.linalg.cholesky(S)) <= 1e-12
    assert np.array_equal(np.triu(A, 1), upper)


def test_la_chol_failures_leave_a_untouched(la):
    A = np.asfortranarray(np.diag([1.0, -1.0, 2.0]))
    before = A.copy()
    assert not la.la_chol(A)
    assert np.array_equal(A, before)
    A[2, 0] = np.nan
    assert not la.la_chol(A)


@pytest.mark.parametrize("n,nrhs", [(70, 1), (300, 37)])
def test_la_trsm_and_trsv(la, n, nrhs):
    rng = np.random.default_rng(n + nrhs)
    L = np.asfortranarray(np.linalg.cholesky(_spd(rng, n)))
    X = np.asfortranarray(rng.standard_normal((n, nrhs)))
    ref = sla.solve_triangular(L, X, lower=True)
    out = la.la_trsm(L, X.copy(order="F"))
    assert _rel(out, ref) <= 1e-12
    x = rng.standard_normal(n)
    assert _rel(la.la_trsv_lower(L, x.copy()), sla.solve_triangular(L, x, lower=True)) <= 1e-12
    assert _rel(la.la_trsv_transpose(L, x.copy()), sla.solve_triangular(L, x, lower=True, trans="T")) <= 1e-12


@pytest.mark.parametrize("n,k,beta", [(130, 200, 0.0), (75, 33, -0.5)])
def test_la_syrk(la, n, k, beta):
    """BLAS syrk('L', 'T'): the lower triangle of S gets alpha P'P + beta S, the upper one is kept."""
    rng = np.random.default_rng(n)
    P = np.asfortranarray(rng.standard_normal((k, n)))
    S = np.asfortranarray(rng.standard_normal((n, n)))
    want = 1.5 * P.T @ P + beta * S
    upper = np.triu(S, 1).copy()
    la.la_syrk(S, P, 1.5, beta)
    assert _rel(np.tril(S), np.tril(want)) <= 1e-12
    assert np.array_equal(np.triu(S, 1), upper)


@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_la_mul_owned(la, ta, tb):
    rng = np.random.default_rng(int(ta) * 2 + int(tb))
    m, n, k = 97, 130, 65
    A = np.asfortranarray(rng.standard_normal((k, m) if ta else (m, k)))
    B = np.asfortranarray(rng.standard_normal((n, k) if tb else (k, n)))
    Cm = np.asfortranarray(rng.standard_normal((m, n)))
    want = 0.75 * (A.T if ta else A) @ (B.T if tb else B) + 2.0 * Cm
    la.la_mul_owned(Cm, A, B, 0.75, 2.0, ta=ta, tb=tb)
    assert _rel(Cm, want) <= 1e-12


def test_la_vector_ops(la):
    rng = np.random.default_rng(9)
    x, y = rng.standard_normal(10007), rng.standard_normal(10007)
    assert abs(la.la_dot(x, y) - float(x @ y)) <= 1e-12 * float(np.abs(x) @ np.abs(y))
    assert la.la_norminf(x) == float(np.max(np.abs(x)))
    assert la.la_norminf(np.zeros(0)) == 0.0 and la.la_dot(np.zeros(0), np.zeros(0)) == 0.0
    Y = y.copy()
    la.la_axpby_owned(2.0, x, -0.5, Y)
    assert _rel(Y, 2.0 * x - 0.5 * y) <= 1e-15
