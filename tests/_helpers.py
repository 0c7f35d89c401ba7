"""Shared fixture data for the oracle and GPU tests."""

from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden" / "dense_fixtures.npz"


def golden():
    return np.load(GOLDEN, allow_pickle=False)


def owner_problem(m=300, blocks=8, side=6):
    """Deterministic dense problem of schur_scheduler_regressions.jl:388-406.

    A_i^(l)[1,1] = (i+l)/17, A_i^(l)[side,side] = (i*l)/23 (1-based), C_l = I,
    c = ones.  At X = Y = I the Schur matrix is the Gram closed form
    S_ij = sum_l (i+l)(j+l)/17^2 + (i l)(j l)/23^2.
    """
    A = np.zeros((blocks, m, side, side))
    ii = np.arange(1, m + 1, dtype=np.float64)
    S = np.zeros((m, m))
    for l in range(1, blocks + 1):
        a = (ii + l) / 17.0
        b = (ii * l) / 23.0
        A[l - 1, :, 0, 0] = a
        A[l - 1, :, side - 1, side - 1] = b
        S += np.outer(a, a) + np.outer(b, b)
    C = [np.eye(side) for _ in range(blocks)]
    return list(A), C, np.ones(m), S


def rel_fro_lower(A, B):
    La, Lb = np.tril(A), np.tril(B)
    return float(np.linalg.norm(La - Lb) / max(np.linalg.norm(Lb), 1e-300))


def random_pd(rng, k, cond=10.0):
    Q, _ = np.linalg.qr(rng.standard_normal((k, k)))
    ev = np.exp(rng.uniform(0, np.log(cond), k))
    return (Q * ev) @ Q.T


def regularization_matrix(m, smallest, seed=9):
    """S = Q diag(smallest, linspace(1,2)[2:]) Q' (kkt_regressions.jl:927-1000)."""
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((m, m)))
    ev = np.linspace(1.0, 2.0, m)
    ev[0] = smallest
    return (Q * ev) @ Q.T


# Refinement cases (kkt.jl:3891-3963): the factor is of S, the refinement runs against
# M = S + rho*G with G = Q diag(ev*u) Q' sharing S's eigenvectors Q, so each correction
# multiplies the error by -rho*u (componentwise in that basis): the contraction per step is
# chosen, not left to rounding.  Each entry: (name, rho, u mode, fixed policy, max steps, expected steps).
REFINE_CASES = [
    ("cap", 0.1, "spread", False, 8, 8),        # >= 2x decrease every step, never reaches 64 eps
    ("slow", 0.7, "ones", False, 8, 1),         # 0.7x decrease: stops after the first correction
    ("diverge", 1.5, "ones", False, 8, 0),      # the correction grows the residual: rolled back
    ("fixed", 0.1, "spread", True, 3, 3),       # fixed policy: exactly refine_steps passes
    ("fixed_diverge", 1.5, "ones", True, 3, 0),  # fixed policy rolls back a worsening pass too
]


def refinement_case(rho, mode, m=200, seed=21):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((m, m)))
    ev = np.exp(rng.uniform(0.0, np.log(100.0), m))
    u = np.ones(m) if mode == "ones" else rng.uniform(0.2, 1.0, m)
    S = (Q * ev) @ Q.T
    M = (Q * (ev * (1.0 + rho * u))) @ Q.T
    r = rng.standard_normal(m)
    return (S + S.T) / 2, (M + M.T) / 2, r
