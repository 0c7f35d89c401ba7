"""Lorentz-cone kernels: oracle pins (CPU) and HIP parity through the C-ABI (GPU).

Reference tests restated: test/soc_native_algebra.jl:1-159.
"""

import math

import numpy as np
import pytest

import soc_oracle as O

from sdpx_mi355x.soc import APPLY_HS_INVERSE, APPLY_W, APPLY_WINV, random_interior


def _cone(rng, q, spread=1.0):
    return list(random_interior(rng, [q], spread))


# ---------------------------------------------------------------- oracle pins (CPU)
def test_nt_roundtrip_reference_vectors():
    # soc_native_algebra.jl:94-118
    ok, w, lam, eta, eta2 = O.nt_scaling([3.0, 0.3, -0.2], [2.5, -0.1, 0.25])
    assert ok
    for q in (2, 3, 5):
        s = [3.0] + [0.15] * (q - 1)
        z = [2.0] + [-0.1] * (q - 1)
        ok, w, lam, eta, _ = O.nt_scaling(s, z)
        assert ok
        src = list(np.linspace(0.1, 0.1 * q, q))
        back = O.apply_winv(w, eta, O.apply_w(w, eta, src))
        np.testing.assert_allclose(back, src, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("q", [2, 3, 4, 7, 16, 65])
def test_nt_defining_equations(q):
    # Nesterov-Todd: W z = W^-1 s = lambda, lambda interior, eta^2 = eta * eta
    rng = np.random.default_rng(q)
    for _ in range(20):
        s, z = _cone(rng, q, 3.0), _cone(rng, q, 0.2)
        ok, w, lam, eta, eta2 = O.nt_scaling(s, z)
        assert ok and eta2 == eta * eta
        scale = max(np.max(np.abs(lam)), 1.0)
        assert np.max(np.abs(np.array(O.apply_w(w, eta, z)) - lam)) <= 1e-12 * scale
        assert np.max(np.abs(np.array(O.apply_winv(w, eta, s)) - lam)) <= 1e-12 * scale
        assert O.is_interior(lam) and O.is_interior(w)
        # (W'W)^-1 = W^-1 W^-1 (W symmetric)
        x = list(rng.standard_normal(q))
        two = O.apply_winv(w, eta, O.apply_winv(w, eta, x))
        np.testing.assert_allclose(O.apply_hs_inverse(w, eta2, x), two, rtol=1e-10, atol=1e-12)


def test_nt_scaling_rejects_non_interior():
    ok, w, lam, eta, eta2 = O.nt_scaling([1.0, 1.0, 0.0], [2.0, 0.1, 0.1])
    assert not ok and w is None and eta == 0.0
    assert not O.nt_scaling([1.0, float("nan")], [1.0, 0.0])[0]


def test_fraction_to_boundary_reference_cases():
    # soc_native_algebra.jl:21-33
    assert math.isclose(O.fraction_to_boundary([1e150, 0.0, 0.0], [-1e150, 1e150, 0.0]), 0.5, rel_tol=1e-12)
    assert O.fraction_to_boundary([1.0, 0.0], [-0.5, 0.5]) == 1.0
    assert math.isclose(O.fraction_to_boundary([1.0, 0.0], [-1.0, 1.0]), 0.5, rel_tol=1e-12)


@pytest.mark.parametrize("q", [2, 3, 6, 20])
def test_fraction_to_boundary_brackets_the_boundary(q):
    rng = np.random.default_rng(100 + q)
    hit = 0
    for _ in range(200):
        s = _cone(rng, q, 0.5)
        ds = list(3.0 * rng.standard_normal(q))
        t = O.fraction_to_boundary(s, ds)
        assert 0.0 <= t <= 1.0
        inside = [a + t * (1 - 1e-9) * b for a, b in zip(s, ds)]
        assert O.is_interior(inside)
        if t < 1.0:
            hit += 1
            outside = [a + t * (1 + 1e-7) * b for a, b in zip(s, ds)]
            assert not O.is_interior(outside)
    assert hit > 20


def test_strict_step_reference_cases():
    # soc_native_algebra.jl:35-63
    assert O.strict_step([[2.0, 0.0]], [[0.0, 0.5]], 1.0, allow_full_step=True) == 1.0
    assert O.strict_step([[2.0, 0.0]], [[0.0, 0.5]], 1.0) == 0.99
    st = O.strict_step([[1.0, 0.0]], [[-0.5, 0.5]], 1.0, allow_full_step=True)
    assert st == 0.99 and O.is_interior([1.0 + st * -0.5, st * 0.5])
    assert O.strict_step([[2.0, 0.0]], [[0.0, 0.5]], 0.5) == 0.495


def test_lorentz_matrix_identities():
    # soc_native_algebra.jl:121-127
    x = [2.0, 0.4, -0.3]
    M = np.array([[x[0] + x[1], x[2]], [x[2], x[0] - x[1]]])
    assert math.isclose(np.linalg.det(M), O.determinant(x), rel_tol=1e-12)
    assert math.isclose(np.linalg.eigvalsh(M)[0], O.margin(x), rel_tol=1e-12)


def test_jordan_solve_inverts_jordan_product():
    rng = np.random.default_rng(7)
    for q in (2, 3, 8):
        left = _cone(rng, q)
        right = list(rng.standard_normal(q))
        x = O.jordan_solve(left, right)
        np.testing.assert_allclose(O.jordan(left, x), right, rtol=1e-11, atol=1e-12)
    assert O.jordan_solve([-2.0, 0.0], [1.0, 0.0]) is None


def test_hkm_metric_closed_form_and_interiority():
    # soc_native_algebra.jl:131-159
    dual = [2.0, 0.25, -0.5]
    ok, d = O.hkm_metric([3.0, 1.0, 0.5], dual)
    assert ok
    ref = [(3 * 2 - 1 * 0.25 + 0.5 * -0.5) / (9 - 1 - 0.25), -(1 * -0.5 + 0.5 * 0.25) / (9 - 1 - 0.25),
           (3 * 2 + 1 * 0.25 - 0.5 * -0.5) / (9 - 1 - 0.25)]
    np.testing.assert_allclose(d, ref, rtol=1e-15)
    assert not O.hkm_metric([-3.0, 1.0, 0.5], dual)[0]
    assert not O.hkm_metric([1.0, 1.0, 0.0], dual)[0]


def _mat(v, half=False):
    a, b, c = v
    if half:
        return np.array([[(a + b) / 2, c / 2], [c / 2, (a - b) / 2]])
    return np.array([[a + b, c], [c, a - b]])


def _coords(Q):
    return [Q[0, 0] + Q[1, 1], Q[0, 0] - Q[1, 1], Q[0, 1] + Q[1, 0]]


@pytest.mark.parametrize("aff", [False, True])
def test_hkm_contractions_match_matrix_algebra(aff):
    rng = np.random.default_rng(11 + aff)
    for _ in range(50):
        x = _cone(rng, 3)
        z = _cone(rng, 3)
        p, ap, ad, dx = (list(rng.standard_normal(3)) for _ in range(4))
        target = float(rng.uniform(0.1, 2.0))
        X, Y, P = _mat(x), _mat(z, half=True), _mat(p)
        R = target * np.eye(2) - X @ Y
        if aff:
            R -= _mat(ap) @ _mat(ad, half=True)
        Xi = np.linalg.inv(X)
        rhs = _coords(Xi @ (P @ Y - R))
        rec = _coords(Xi @ (R - _mat(dx) @ Y))
        np.testing.assert_allclose(O.hkm_rhs_coordinates(x, z, p, ap, ad, target, aff), rhs, rtol=1e-9, atol=1e-10)
        np.testing.assert_allclose(O.hkm_recovery(x, z, dx, ap, ad, target, aff), rec, rtol=1e-9, atol=1e-10)


def test_planted_sign_error_is_caught(monkeypatch):
    orig = O.apply_winv

    def bad(w, eta, src):
        out = orig(w, eta, src)
        out[0] = -out[0]
        return out

    monkeypatch.setattr(O, "apply_winv", bad)
    with pytest.raises(AssertionError):
        test_nt_roundtrip_reference_vectors()


# ---------------------------------------------------------------- GPU parity
def _ulps(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    ia = a.view(np.int64).astype(np.float64)
    ib = b.view(np.int64).astype(np.float64)
    same_sign = np.sign(a) == np.sign(b)
    d = np.where(same_sign, np.abs(ia - ib), np.where((a == 0) & (b == 0), 0, np.inf))
    return float(np.max(d)) if d.size else 0.0


def _mixed_dims(rng, n):
    choices = np.array([1, 2, 3, 4, 5, 8, 17, 64])
    return rng.choice(choices, size=n)


@pytest.fixture
def cones(gpu_lib):
    from sdpx_mi355x.soc import LorentzCones

    made = []

    def make(dims):
        c = LorentzCones(dims)
        made.append(c)
        return c

    yield make
    for c in made:
        c.close()


def _split(v, dims):
    return np.split(v, np.cumsum(dims)[:-1])


@pytest.mark.gpu
@pytest.mark.parametrize("q", [None, 2, 3, 4, 5, 8])
def test_gpu_nt_scaling_and_apply_match_oracle(cones, q):
    # q None: mixed dimensions (offset-table kernels); q fixed: every cone of
    # dimension q (register-resident kernels)
    rng = np.random.default_rng(1 if q is None else 10 + q)
    if q is None:
        dims = _mixed_dims(rng, 1500)
        dims[dims == 1] = 2
    else:
        dims = np.full(1500, q)
    s = random_interior(rng, dims, 2.0)
    z = random_interior(rng, dims, 0.3)
    # a few cones outside the interior: scaling must fail there and leave w / lambda untouched
    starts = np.concatenate([[0], np.cumsum(dims)[:-1]])
    bad = rng.choice(len(dims), 20, replace=False)
    s[starts[bad]] = -1.0
    c = cones(dims)
    c.set(0, s)
    c.set(1, z)
    sentinel = np.full(c.total, 7.0)
    c.set(2, sentinel)
    c.set(3, sentinel)
    nfail = c.nt_scaling(0, 1, 2, 3)
    w, lam = c.get(2), c.get(3)
    eta, eta2, ok = c.scaling()
    assert nfail == len(bad)
    worst = 0.0
    for i, (si, zi) in enumerate(zip(_split(s, dims), _split(z, dims))):
        rok, rw, rl, re, re2 = O.nt_scaling(list(si), list(zi))
        assert ok[i] == rok
        sl = slice(starts[i], starts[i] + dims[i])
        if rok:
            worst = max(worst, _ulps(w[sl], rw), _ulps(lam[sl], rl), _ulps([eta[i], eta2[i]], [re, re2]))
        else:
            assert np.all(w[sl] == 7.0) and np.all(lam[sl] == 7.0) and eta[i] == 0.0
    assert worst <= 2, worst
    # W, W^-1 and (W'W)^-1 on a fresh vector, one with dst aliasing src
    x = rng.standard_normal(c.total)
    c.set(4, x)
    for op, fn in ((APPLY_W, O.apply_w), (APPLY_WINV, O.apply_winv), (APPLY_HS_INVERSE, O.apply_hs_inverse)):
        c.set(5, x)
        c.apply(op, 2, 4, 6)
        c.apply(op, 2, 5, 5)
        y, y_alias = c.get(6), c.get(5)
        assert np.array_equal(y, y_alias)
        worst = 0.0
        for i in np.flatnonzero(ok):
            sl = slice(starts[i], starts[i] + dims[i])
            scal = eta2[i] if op == APPLY_HS_INVERSE else eta[i]
            worst = max(worst, _ulps(y[sl], fn(list(w[sl]), scal, list(x[sl]))))
        assert worst <= 2, (op, worst)


@pytest.mark.gpu
def test_gpu_jordan_solve_inverse_margins(cones):
    rng = np.random.default_rng(2)
    dims = _mixed_dims(rng, 1000)
    dims[dims == 1] = 2
    starts = np.concatenate([[0], np.cumsum(dims)[:-1]])
    left = random_interior(rng, dims, 1.0)
    left[starts[5]] = -3.0  # one divisor outside the cone
    right = rng.standard_normal(left.size)
    c = cones(dims)
    c.set(0, left)
    c.set(1, right)
    c.set(2, np.full(left.size, 5.0))
    c.jordan(0, 1, 3)
    assert c.jordan_solve(0, 1, 2) == 1
    c.set(4, np.full(left.size, 5.0))
    assert c.inverse(0, 4) == 1
    mg, nint = c.margins(0)
    xs, prods, invs = c.get(2), c.get(3), c.get(4)
    worst = 0.0
    for i in range(len(dims)):
        sl = slice(starts[i], starts[i] + dims[i])
        l, r = list(left[sl]), list(right[sl])
        worst = max(worst, _ulps(prods[sl], O.jordan(l, r)), _ulps([mg[i]], [O.margin(l)]))
        ref = O.jordan_solve(l, r)
        if ref is None:
            assert i == 5 and np.all(xs[sl] == 5.0) and np.all(invs[sl] == 5.0)
        else:
            det = O.determinant(l)
            worst = max(worst, _ulps(xs[sl], ref), _ulps(invs[sl], [l[0] / det] + [-v / det for v in l[1:]]))
    assert worst <= 2, worst
    assert nint == len(dims) - 1


@pytest.mark.gpu
def test_gpu_fraction_to_boundary_and_strict_step(cones):
    rng = np.random.default_rng(3)
    dims = _mixed_dims(rng, 2000)
    starts = np.concatenate([[0], np.cumsum(dims)[:-1]])
    s = random_interior(rng, dims, 0.5)
    ds = 2.0 * rng.standard_normal(s.size)
    c = cones(dims)
    c.set(0, s)
    c.set(1, ds)
    mn, per = c.fraction_to_boundary(0, 1, per_cone=True)
    ref = [O.fraction_to_boundary(list(s[b:b + q]), list(ds[b:b + q])) for b, q in zip(starts, dims)]
    assert _ulps(per, ref) <= 2
    assert mn == min(ref)
    st = [list(s[b:b + q]) for b, q in zip(starts, dims)]
    dr = [list(ds[b:b + q]) for b, q in zip(starts, dims)]
    for bound, full in ((mn, False), (1.0, True), (0.5, False)):
        assert c.strict_step(0, 1, bound, full) == O.strict_step(st, dr, bound, full)


@pytest.mark.gpu
def test_gpu_strict_step_reference_cases(cones):
    c = cones([2])
    c.set(0, [2.0, 0.0])
    c.set(1, [0.0, 0.5])
    assert c.strict_step(0, 1, 1.0, True) == 1.0
    assert c.strict_step(0, 1, 1.0, False) == 0.99
    assert c.strict_step(0, 1, 0.5, False) == 0.495
    c.set(0, [1.0, 0.0])
    c.set(1, [-0.5, 0.5])
    assert c.strict_step(0, 1, 1.0, True) == 0.99
    c3 = cones([3])
    c3.set(0, [1e150, 0.0, 0.0])
    c3.set(1, [-1e150, 1e150, 0.0])
    assert math.isclose(c3.fraction_to_boundary(0, 1), 0.5, rel_tol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("aff", [False, True])
def test_gpu_fixed_trace_q3_kernels(cones, aff):
    rng = np.random.default_rng(4 + aff)
    n = 3000
    dims = np.full(n, 3)
    x = random_interior(rng, dims, 1.0)
    x[3 * 7] = -x[3 * 7]  # reflected head: metric must refuse it
    z = random_interior(rng, dims, 1.0)
    p, ap, ad = (rng.standard_normal(3 * n) for _ in range(3))
    c = cones(dims)
    for slot, v in enumerate((x, z, p, ap, ad)):
        c.set(slot, v)
    c.set(5, np.full(3 * n, 9.0))
    assert c.hkm_metric(0, 1, 5) == 1
    c.hkm_rhs_coordinates(0, 1, 2, 3, 4, 0.7, aff, 6)
    c.hkm_recovery(0, 1, 2, 3, 4, 0.7, aff, 7)
    met, rhs, rec = c.get(5), c.get(6), c.get(7)
    worst = 0.0
    for i in range(n):
        sl = slice(3 * i, 3 * i + 3)
        args = (list(x[sl]), list(z[sl]))
        ok, d = O.hkm_metric(*args)
        if ok:
            worst = max(worst, _ulps(met[sl], d))
        else:
            assert i == 7 and np.all(met[sl] == 9.0)
        if i != 7:
            worst = max(worst, _ulps(rhs[sl], O.hkm_rhs_coordinates(*args, list(p[sl]), list(ap[sl]), list(ad[sl]),
                                                                   0.7, aff)))
            worst = max(worst, _ulps(rec[sl], O.hkm_recovery(*args, list(p[sl]), list(ap[sl]), list(ad[sl]), 0.7,
                                                            aff)))
    assert worst <= 2, worst


@pytest.mark.gpu
def test_gpu_full_size_properties(cones):
    # 2M cones of dimension 4 (the ladder's soc_many shape, scaled up): NT
    # equations and the W / W^-1 round trip hold everywhere
    rng = np.random.default_rng(5)
    n, q = 2_000_000, 4
    dims = np.full(n, q)
    s = random_interior(rng, dims, 2.0)
    z = random_interior(rng, dims, 0.5)
    c = cones(dims)
    c.set(0, s)
    c.set(1, z)
    assert c.nt_scaling(0, 1, 2, 3) == 0
    c.apply(APPLY_W, 2, 1, 4)       # W z
    c.apply(APPLY_WINV, 2, 0, 5)    # W^-1 s
    lam, wz, wis = c.get(3), c.get(4), c.get(5)
    scale = np.maximum(np.abs(lam).reshape(n, q).max(axis=1), 1.0).repeat(q)
    assert np.max(np.abs(wz - lam) / scale) < 1e-12
    assert np.max(np.abs(wis - lam) / scale) < 1e-12
    x = rng.standard_normal(n * q)
    c.set(6, x)
    c.apply(APPLY_W, 2, 6, 7)
    c.apply(APPLY_WINV, 2, 7, 7)
    np.testing.assert_allclose(c.get(7), x, rtol=1e-9, atol=1e-9)


# ---------------------------------------------------------------- metric assembly (soc_native.jl:1119-1218)
@pytest.mark.parametrize("q", [2, 3, 4, 7])
def test_metric_matrix_maps_s_to_z(q):
    """The NT scaling is defined by W z = W^-1 s = lambda, so (W'W)^-1 s = z (W symmetric): the oracle's
    metric block is pinned by the NT equations, and it is symmetric positive definite."""
    rng = np.random.default_rng(q)
    s, z = _cone(rng, q, 1.5), _cone(rng, q, 0.7)
    ok, w, lam, eta, eta2 = O.nt_scaling(s, z)
    assert ok
    M = O.metric_matrix(w, eta2)
    assert np.allclose(M @ np.array(s), z, rtol=1e-13, atol=1e-13 * np.max(np.abs(z)))
    assert np.allclose(M, M.T, rtol=0, atol=1e-14 * np.max(np.abs(M)))
    assert np.min(np.linalg.eigvalsh((M + M.T) / 2)) > 0


def test_add_metric_block_structure():
    """Cones whose coefficient rows select disjoint variables give a block-diagonal Hessian whose
    blocks are the metric matrices (each pinned by the NT equations above)."""
    rng = np.random.default_rng(3)
    dims = [3, 2, 4]
    n = sum(dims)
    H = np.zeros((n, n))
    blocks, off = [], 0
    for q in dims:
        s, z = _cone(rng, q, 1.2), _cone(rng, q, 0.9)
        ok, w, _, _, eta2 = O.nt_scaling(s, z)
        A = np.zeros((q, n))
        A[:, off:off + q] = np.eye(q)
        O.add_metric(H, A, w, eta2)
        blocks.append((off, q, O.metric_matrix(w, eta2)))
        off += q
    for o, q, M in blocks:  # 0/1 selections: every product is exact
        assert np.array_equal(H[o:o + q, o:o + q], M)
    mask = np.ones((n, n), bool)
    for o, q, _ in blocks:
        mask[o:o + q, o:o + q] = False
    assert np.all(H[mask] == 0.0)


@pytest.mark.gpu
@pytest.mark.parametrize("dims_kind", ["mixed", "uniform4"])
def test_gpu_metric_matches_oracle(gpu_lib, dims_kind):
    """sdpx_soc_metric (one W^-1 pass over the coefficient slab, then the FP64 Gram kernel) against the
    oracle's per-cone dense assembly, accumulating onto H0; a cone whose NT scaling failed is reported."""
    from sdpx_mi355x.soc import LorentzCones

    rng = np.random.default_rng(17)
    ncones, n = 700, 333
    dims = rng.choice([2, 3, 4, 5, 8, 13], size=ncones) if dims_kind == "mixed" else np.full(ncones, 4)
    s = random_interior(rng, dims, 1.5)
    z = random_interior(rng, dims, 0.7)
    A = rng.standard_normal((int(dims.sum()), n)) / np.sqrt(n)
    H0 = rng.standard_normal((n, n))
    H0 = H0 + H0.T
    with LorentzCones(dims) as c:
        c.set(0, s)
        c.set(1, z)
        c.set_constraints(A)
        assert c.metric(2)[1] == ncones  # before any NT scaling every cone is unscaled
        assert c.nt_scaling(0, 1, 2, 3) == 0
        H, nf = c.metric(2, H0=H0)
        assert nf == 0
        eta, eta2, ok = c.scaling()
        w = c.get(2)
        ref = H0.copy()
        off = 0
        for ci, q in enumerate(dims):
            O.add_metric(ref, A[off:off + q], list(w[off:off + q]), float(eta2[ci]))
            off += q
        rel = np.linalg.norm(H - ref) / np.linalg.norm(ref - H0)
        print(f"{dims_kind}: |H - oracle| / |sum A'MA| = {rel:.2e}", flush=True)
        assert rel <= 1e-13
        assert np.array_equal(H, H.T)
        # a cone outside the interior: its NT scaling fails and the metric reports it
        s_bad = s.copy()
        s_bad[0] = -1.0
        c.set(0, s_bad)
        assert c.nt_scaling(0, 1, 2, 3) == 1
        H2, nf = c.metric(2, download=True)
        assert nf == 1 and H2 is not None
