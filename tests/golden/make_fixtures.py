"""Write the golden fixtures of the dense Newton-step path to ``dense_fixtures.npz``.

The inputs restate data held by SDPX.jl's own tests; the expected outputs are
computed here by brute force in 50-digit mpmath (independent of the
triangular-solve / Gram algorithm under test) or by closed forms:

* ``t1_*``  -- test/correctness.jl:71-94 (T1 data), 230-243 (iteration pin 19)
* ``kkt2_*`` -- test/kkt_regressions.jl:488-565 (two 2x2 blocks, 4 variables,
  explicit X and Y); expected S_ij = sum_l tr(A_i X_l^-1 A_j Y_l)
* ``tiny{n}_*`` -- seeded random tiny SDPs (k <= 5, m <= 6) with random PD X, Y;
  expected S by brute force.

Run: ``python tests/golden/make_fixtures.py``.
"""

from pathlib import Path

import mpmath
import numpy as np

mpmath.mp.dps = 50
OUT = Path(__file__).resolve().parent / "dense_fixtures.npz"


def brute_schur(A_blocks, X_blocks, Y_blocks):
    """S_ij = sum_l tr(A_i^l X_l^-1 A_j^l Y_l) in 50-digit arithmetic."""
    m = A_blocks[0].shape[0]
    S = mpmath.zeros(m, m)
    for A, X, Y in zip(A_blocks, X_blocks, Y_blocks):
        Xi = mpmath.matrix(X.tolist()) ** -1
        Ym = mpmath.matrix(Y.tolist())
        Am = [mpmath.matrix(A[i].tolist()) for i in range(m)]
        left = [Am[i] * Xi for i in range(m)]
        right = [Am[j] * Ym for j in range(m)]
        for i in range(m):
            for j in range(i + 1):
                M = left[i] * right[j]
                v = sum(M[q, q] for q in range(M.rows))
                S[i, j] += v
                if i != j:
                    S[j, i] += v
    return np.array([[float(S[i, j]) for j in range(m)] for i in range(m)])


def main():
    fx = {}
    # T1
    A = np.zeros((2, 2, 2))
    A[0, 0, 0] = 1.0
    A[1, 1, 1] = 1.0
    fx["t1_A"] = A
    fx["t1_C"] = np.array([[0.0, 1.0], [1.0, 0.0]])
    fx["t1_c"] = np.array([2.0, 3.0])
    fx["t1_popt"] = np.array(2 * np.sqrt(6.0))
    fx["t1_iterations_fixed"] = np.array(19)

    # kkt_regressions.jl two-block fixture
    c1 = np.zeros((4, 2, 2))
    c2 = np.zeros((4, 2, 2))
    c1[0] = [[0.0, 1.0], [1.0, 0.0]]
    c1[1] = [[1.0, 0.0], [0.0, -1.0]]
    c2[2] = [[0.0, 1.0], [1.0, 0.0]]
    c2[3] = [[1.0, 0.0], [0.0, -1.0]]
    X = [np.array([[2.0, 0.1], [0.1, 1.5]]), np.array([[1.8, -0.2], [-0.2, 2.2]])]
    Y = [np.array([[1.4, 0.2], [0.2, 1.9]]), np.array([[2.1, -0.1], [-0.1, 1.6]])]
    fx["kkt2_A1"], fx["kkt2_A2"] = c1, c2
    fx["kkt2_X1"], fx["kkt2_X2"] = X
    fx["kkt2_Y1"], fx["kkt2_Y2"] = Y
    fx["kkt2_S"] = brute_schur([c1, c2], X, Y)

    # tiny random problems, brute force
    rng = np.random.default_rng(20240601)
    for n, (k, mm) in enumerate([(2, 3), (3, 5), (4, 6), (5, 4)]):
        Ar = rng.standard_normal((mm, k, k))
        Ar = (Ar + Ar.transpose(0, 2, 1)) / 2
        G1 = rng.standard_normal((k, k))
        G2 = rng.standard_normal((k, k))
        Xr = G1 @ G1.T + k * np.eye(k)
        Yr = G2 @ G2.T + 0.5 * np.eye(k)
        fx[f"tiny{n}_A"] = Ar
        fx[f"tiny{n}_X"] = Xr
        fx[f"tiny{n}_Y"] = Yr
        fx[f"tiny{n}_S"] = brute_schur([Ar], [Xr], [Yr])
    np.savez_compressed(OUT, **fx)
    print(f"wrote {OUT} ({len(fx)} arrays)")


if __name__ == "__main__":
    main()
