"""CPU FP64 restatement of the implicit-differentiation adjoint — TEST INFRASTRUCTURE ONLY.

The reference has no differentiation (no ChainRulesCore dependency in
Project.toml), so this module is not a restatement of a reference function:
it states the implicit-differentiation formulas the HIP adjoint implements,
on top of the oracle's own Schur assembly (sdpx_oracle.schur_build,
schur.jl:511-547), and the central-path map whose finite differences pin
them.  Parity with the reference is unpinned by construction; the adjoint is
pinned by central finite differences of ``central_point`` (tests/test_adjoint.py).

Core form (sdpx_oracle.Problem):  sum_i x_i A_i - X - C = 0,  <A_i, Y> = c_i,
XY = mu I.  The user's dual-form model  max b'y  s.t.  C_std - sum y_i A_std,i >= 0
maps to it by A = -A_std, C = -C_std, c = -b, x = y; its primal matrix X* is
the core Y.  Loss L = <G, X*> + h'y* = <G, Y> + h'x.

Linearising the three conditions (dY = -sym(X^-1 dX Y), HKM) gives
S dx = [<A_i, X^-1 dC Y>] - dc with the Schur matrix S_ij = <A_i, X^-1 A_j Y>;
the adjoint of that map is
  W = sym(X^-1 G Y),  lambda = S^-1 (h - [<A_i, W>]),
  dL/dC = sym(X^-1 A_lambda Y) + W,  dL/dc = -lambda      (core)
  dL/dC_std = -dL/dC,  dL/db = lambda                       (user's model).
"""

from __future__ import annotations

import numpy as np
import scipy.linalg as sla

import sdpx_oracle as O


def _sym(M):
    return (M + M.T) / 2.0


def schur_at(prob: O.Problem, X, Y) -> np.ndarray:
    LX, okx = O.kchol(X)
    MY, oky = O.kchol(Y)
    if not (okx and oky):
        raise ValueError("iterate not positive definite")
    return O.schur_build(prob, [LX], [MY])


def adjoint(prob: O.Problem, x, X, Y, G, h):
    """(dL/dC_std, dL/db) at the iterate (single block)."""
    S = schur_at(prob, X, Y)
    Xi = np.linalg.inv(X)
    W = _sym(Xi @ G @ Y)
    g = np.asarray(h, dtype=np.float64) - O.contract(prob, 0, W)
    lam = sla.solve(S, g, assume_a="pos")
    Al = O.build_P(prob, 0, lam)
    dC_core = _sym(Xi @ Al @ Y) + W
    return -dC_core, lam


def center(prob: O.Problem, x, X, Y, mu: float, iterations: int = 6):
    """Newton steps on the central-path equations with mu fixed (full steps, 0.9 backtracking if needed)."""
    x, X, Y = np.array(x, dtype=np.float64), np.array(X, dtype=np.float64), np.array(Y, dtype=np.float64)
    k = X.shape[0]
    for _ in range(iterations):
        P = O.build_P(prob, 0, x) - X - prob.C[0]
        d = prob.c - O.contract(prob, 0, Y)
        R = mu * np.eye(k) - X @ Y
        S = schur_at(prob, X, Y)
        Xi = np.linalg.inv(X)
        Z = Xi @ (P @ Y - R)
        r = -(d + O.contract(prob, 0, Z))
        dx = sla.solve(S, r, assume_a="pos")
        dX = O.build_P(prob, 0, dx) + P
        dY = _sym(Xi @ (R - dX @ Y))
        t = 1.0
        while not (O.kchol(X + t * dX)[1] and O.kchol(Y + t * dY)[1]):
            t *= 0.9
        x, X, Y = x + t * dx, X + t * dX, Y + t * dY
    return x, X, Y


def loss(x, Y, G, h) -> float:
    return float(np.sum(G * Y) + np.dot(h, x))


def perturbed(prob: O.Problem, dC_std=None, db=None) -> O.Problem:
    """The same problem with C_std + dC_std and b + db (core: C - dC_std, c - db)."""
    C = prob.C[0] - (dC_std if dC_std is not None else 0.0)
    c = prob.c - (db if db is not None else 0.0)
    A = prob.Av[0].T.reshape(prob.m, prob.k[0], prob.k[0])
    return O.ingest(c, [A], [C])


def central_point_loss(prob: O.Problem, x, X, Y, mu: float, G, h, iterations: int = 6) -> float:
    """L at the central-path point of mu, reached by centering from (x, X, Y)."""
    xc, _, Yc = center(prob, x, X, Y, mu, iterations)
    return loss(xc, Yc, G, h)


def optimum_adjoint(prob: O.Problem, x, X, Y, G, h, rank_tol: float = 1e-5):
    """(dL/dC_std, dL/db) of the optimum itself (single block), by the implicit function theorem on
    the optimality conditions A(x) - X - C = 0, <A_i, Y> = c_i, XY = 0 at a strictly complementary,
    nondegenerate solution.

    With Q = [V U] the eigenvectors of X (V: its range, rank r = eigenvalues above rank_tol times the
    largest; U: its null space, where Y lives), Lx = diag of X on V and Yu = U'YU, the linearised
    complementarity reads dX_UU = 0, dY_VV = 0, Lx dY_VU + dX_VU Yu = 0.  Eliminating dY_VU leaves a
    square system in (dx, dY_UU):
        sum_i dx_i (A_i)_UU = dC_UU,
        -H dx + [<(A_i)_UU, dY_UU>] = dc - 2 <(A_i)_VU, Lx^-1 dC_VU Yu>,
    H_ij = 2 <(A_i)_VU, Lx^-1 (A_j)_VU Yu>; its adjoint gives the gradient of L = <G, Y> + h'x.
    Returns None when the rank split is not clean (no gap in X's spectrum or Y not PD on U)."""
    k, m = X.shape[0], prob.m
    ev, Q = np.linalg.eigh(_sym(X))
    order = np.argsort(ev)[::-1]
    ev, Q = ev[order], Q[:, order]
    r = int(np.sum(ev > rank_tol * max(ev[0], 1e-300)))
    if r == 0 or r == k:
        return None
    V, U = Q[:, :r], Q[:, r:]
    lx = ev[:r]
    Yu = _sym(U.T @ Y @ U)
    if np.min(np.linalg.eigvalsh(Yu)) <= rank_tol * np.max(np.abs(Yu)):
        return None
    A = prob.Av[0].T.reshape(m, k, k)  # A[i] = A_i (symmetric)
    At = np.einsum("ap,iab,bq->ipq", Q, A, Q)
    Bvu, Buu = At[:, :r, r:], At[:, r:, r:]
    ku = k - r
    iu, ju = np.triu_indices(ku)
    t = len(iu)
    w = np.where(iu == ju, 1.0, 2.0)
    Buu_mat = Buu[:, iu, ju].T                               # t x m
    scaled = Bvu / lx[None, :, None] @ Yu                    # Lx^-1 B_j Yu
    H = 2.0 * np.einsum("iab,jab->ij", Bvu, scaled)
    M = np.zeros((t + m, m + t))
    M[:t, :m] = Buu_mat
    M[t:, :m] = -H
    M[t:, m:] = (Buu_mat * w[:, None]).T
    Gt = Q.T @ _sym(np.asarray(G, dtype=np.float64)) @ Q
    Gvu, Guu = Gt[:r, r:], Gt[r:, r:]
    K = (Gvu / lx[:, None]) @ Yu                             # Lx^-1 G_VU Yu
    lx_vec = np.asarray(h, dtype=np.float64) - 2.0 * np.einsum("ab,iab->i", Gvu, scaled)
    lu_vec = w * Guu[iu, ju]
    nu = np.linalg.solve(M.T, np.concatenate([lx_vec, lu_vec]))
    nu_u, nu_x = nu[:t], nu[t:]
    Ghat = np.zeros((k, k))
    Guu_hat = np.zeros((ku, ku))
    Guu_hat[iu, ju] = np.where(iu == ju, nu_u, nu_u / 2.0)
    Guu_hat = Guu_hat + np.triu(Guu_hat, 1).T
    Gvu_hat = -(np.einsum("i,iab->ab", nu_x, Bvu) / lx[:, None]) @ Yu + K
    Ghat[r:, r:] = Guu_hat
    Ghat[:r, r:] = Gvu_hat
    Ghat[r:, :r] = Gvu_hat.T
    dC_core = Q @ Ghat @ Q.T
    # core dL/dc = nu_x; the user's model: dL/dC_std = -dL/dC, dL/db = -dL/dc
    return -dC_core, -nu_x


def extrapolated_adjoint(prob: O.Problem, x, X, Y, G, h, mu: float, ratio: float = 0.1, iterations: int = 12):
    """The central-path adjoint at mu and at ratio * mu, combined by one Richardson step for a
    derivative linear in mu near mu = 0: D(0) ~ (D(ratio mu) - ratio D(mu)) / (1 - ratio)."""
    x1, X1, Y1 = center(prob, x, X, Y, mu, iterations)
    D1 = adjoint(prob, x1, X1, Y1, G, h)
    x2, X2, Y2 = center(prob, x1, X1, Y1, ratio * mu, iterations)
    D2 = adjoint(prob, x2, X2, Y2, G, h)
    return tuple((b - ratio * a) / (1.0 - ratio) for a, b in zip(D1, D2))
