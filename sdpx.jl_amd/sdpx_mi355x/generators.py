"""Seeded synthetic SDP generators for the dense and sparse workloads.

All generators build the standard-form problem

    min <C, X>  s.t.  <A_i, X> = b_i,  X >= 0

with a strictly feasible primal-dual start X0 = I, y0 ~ N(0,1), Z0 = I, via
b_i = tr(A_i) and C = sum_i y0_i A_i + I, and return it in SDPX's core
geometric form (the lowering of the dual-form model, ir/lower_sdp.jl:244-300):

    core x = y,  A_i^core = -A_i,  C^core = -C,  c = -b,
    start: x0 = y0, X0 = Z0 = I, Y0 = X0 = I.

Layouts follow the reference's dense panel (types/constraints.jl:26-28): the
per-block panel ``Av`` is k^2 x m column-major with column i = vec(A_i).  Here
it is returned as a C-contiguous (m, k*k) array whose row i is vec(A_i), which
is the same bytes.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class DenseSDP:
    """One core-form SDP with a single dense PSD block."""

    k: int
    m: int
    Av_rows: np.ndarray  # (m, k*k) float64, row i = vec(A_i^core)
    C: np.ndarray        # (k, k) core C
    c: np.ndarray        # (m,) core c = -b
    x0: np.ndarray       # (m,) = y0
    X0: np.ndarray       # (k, k) = I
    Y0: np.ndarray       # (k, k) = I

    def A_blocks(self) -> np.ndarray:
        """Core A as an (m, k, k) array (the layout SDPX.ingest accepts)."""
        return self.Av_rows.reshape(self.m, self.k, self.k)


def _symmetric_gaussian(rng: np.random.Generator, m: int, k: int, scale: float) -> np.ndarray:
    """m symmetric k x k matrices: upper-triangle iid N(0,1)*scale, mirrored."""
    iu = np.triu_indices(k)
    out = np.zeros((m, k, k))
    vals = rng.standard_normal((m, iu[0].size)) * scale
    out[:, iu[0], iu[1]] = vals
    out[:, iu[1], iu[0]] = vals
    return out


def dense_feasible(k: int, m: int, seed: int = 0) -> DenseSDP:
    """Dense Gaussian generator of the plain FP64 workloads (k=n PSD block, m constraints).

    A_i symmetric with upper-triangle entries iid N(0,1)/sqrt(n), mirrored;
    X0 = I, b_i = tr(A_i), y0 ~ N(0,1), C = sum y0_i A_i + I.
    """
    rng = np.random.default_rng(seed)
    # iid N(0,1)/sqrt(k) on and above the diagonal, mirrored below; built in
    # one (m, k, k) buffer in chunks so multi-GB panels need no extra copies
    A = np.empty((m, k, k))
    scale = 1.0 / np.sqrt(k)
    lower = np.tril_indices(k, -1)
    chunk = max(1, (1 << 27) // (k * k))
    for i0 in range(0, m, chunk):
        blk = A[i0:i0 + chunk]
        rng.standard_normal(out=blk)
        blk *= scale
        blk[:, lower[0], lower[1]] = blk[:, lower[1], lower[0]]
    y0 = rng.standard_normal(m)
    b = np.trace(A, axis1=1, axis2=2).copy()
    Cstd = (y0 @ A.reshape(m, k * k)).reshape(k, k) + np.eye(k)
    np.negative(A, out=A)  # core sign: A_core = -A
    return DenseSDP(
        k=k,
        m=m,
        Av_rows=A.reshape(m, k * k),
        C=-Cstd,
        c=-b,
        x0=y0.copy(),
        X0=np.eye(k),
        Y0=np.eye(k),
    )


@dataclass
class SparseSDP:
    """One core-form SDP whose A_i are sparse (COO of the upper triangle)."""

    k: int
    m: int
    rows: list   # per constraint: int32 row indices (upper triangle, row <= col)
    cols: list
    vals: list
    C: np.ndarray
    c: np.ndarray
    x0: np.ndarray

    def dense_rows(self) -> np.ndarray:
        out = np.zeros((self.m, self.k, self.k))
        for i in range(self.m):
            out[i, self.rows[i], self.cols[i]] = self.vals[i]
            out[i, self.cols[i], self.rows[i]] = self.vals[i]
        return out.reshape(self.m, self.k * self.k)


def sparse_feasible(k: int, m: int, seed: int) -> SparseSDP:
    """Sparse generator of the batched workload: 3k nonzero positions per A_i.

    Positions are uniform over the upper triangle (distinct), values N(0,1);
    the mirrored matrix is symmetric. Feasible construction as in the dense
    generator.
    """
    rng = np.random.default_rng(seed)
    iu = np.triu_indices(k)
    npos = iu[0].size
    nnz = min(3 * k, npos)
    y0 = rng.standard_normal(m)
    # nnz distinct positions per constraint: the first nnz of a random order
    pick = np.argsort(rng.random((m, npos)), axis=1)[:, :nnz]
    r = iu[0][pick].astype(np.int32)
    cc = iu[1][pick].astype(np.int32)
    v = rng.standard_normal((m, nnz))
    diag = r == cc
    b = np.where(diag, v, 0.0).sum(axis=1)
    Cstd = np.eye(k)
    w = y0[:, None] * v
    np.add.at(Cstd, (r.ravel(), cc.ravel()), w.ravel())
    off = ~diag
    np.add.at(Cstd, (cc[off], r[off]), w[off])
    return SparseSDP(k=k, m=m, rows=list(r), cols=list(cc), vals=list(-v), C=-Cstd, c=-b, x0=y0)


def batched_sparse(B: int, seed: int = 0, kmin: int = 8, kmax: int = 64, only=None) -> list:
    """The uneven batched workload: B problems, k ~ U{kmin..kmax}, m ~ U{2k..floor(k(k+1)/4)},
    3k nonzero positions per A_i (sparse_feasible), one seed per problem.

    only: optional member indices; then just those members are built (in that order), from the
    same draws as the whole batch."""
    rng = np.random.default_rng(seed)
    ks = rng.integers(kmin, kmax + 1, size=B)
    shapes = []
    for k in ks:
        k = int(k)
        mlo, mhi = 2 * k, max(2 * k, (k * (k + 1)) // 4)
        shapes.append((k, int(rng.integers(mlo, mhi + 1))))
    members = range(B) if only is None else [int(p) for p in only]
    return [sparse_feasible(*shapes[p], seed=seed * 1_000_003 + p + 1) for p in members]


def batched_dense_layer(B: int, seed: int = 0, k: int = 32, m: int = 128, noise: float = 0.1) -> list:
    """The differentiable-layer workload: B dense Gaussian problems (dense_feasible), with C and b
    perturbed by seeded N(0, noise^2) noise (C's noise symmetric)."""
    rng = np.random.default_rng(seed)
    out = []
    for p in range(B):
        g = dense_feasible(k, m, seed=seed * 1_000_003 + p + 1)
        E = rng.standard_normal((k, k)) * noise
        E = np.triu(E) + np.triu(E, 1).T
        db = rng.standard_normal(m) * noise
        # core signs: C_core = -C, c = -b
        out.append(DenseSDP(k=k, m=m, Av_rows=g.Av_rows, C=g.C - E, c=g.c - db, x0=g.x0, X0=g.X0, Y0=g.Y0))
    return out
