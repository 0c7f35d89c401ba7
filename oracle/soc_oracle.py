"""CPU oracle for SDPX's Lorentz (second-order) cone kernels — TEST INFRASTRUCTURE ONLY.

Scalar Python restatement of the reference's Lorentz-coordinate kernels
(src/soc_lorentz_kernels.jl, src/soc.jl, src/soc_native.jl) in IEEE Float64,
one operation at a time in the reference's order, so that the HIP kernels
(compiled without floating-point contraction) can be compared bit for bit.
Only ``tests/``, ``__graft_entry__.smoke`` and ``bench.py``'s ``cpu_baseline``
leg may import this module.

A cone vector is v = (v[0], v[1:]) = (head, tail), Q_q = {(t, u): t >= ||u||}.

Parity pins (tests/test_soc.py):

* ``nt_scaling``/``apply_w``/``apply_winv`` -- the reference's vectors and
  round-trip bound (test/soc_native_algebra.jl:94-118) and the NT defining
  equations W z = W^-1 s = lambda (Nesterov-Todd); ``apply_hs_inverse`` --
  (W'W)^-1 = W^-1 W^-1.
* ``fraction_to_boundary`` -- the reference's scaled boundary cases
  (soc_native_algebra.jl:21-33) and brute-force interiority on both sides of
  the root.
* ``strict_step`` -- the reference's policy cases (soc_native_algebra.jl:35-63).
* ``determinant``/``margin`` -- the 2x2 matrix image identities
  (soc_native_algebra.jl:121-127): det and eigmin of [[x0+x1, x2], [x2, x0-x1]].
* ``jordan_solve`` -- left o result = right with ``jordan``; ``jordan`` by its
  bilinear definition.
* ``hkm_metric`` -- the closed-form vector and interiority contract of
  soc_native_algebra.jl:131-159; ``hkm_rhs_coordinates``/``hkm_recovery`` --
  the explicit 2x2 matrix algebra they contract (X^-1 (P Y - R) and
  X^-1 (R - dX Y), Y = Z/2).
"""

from __future__ import annotations

import math

import numpy as np


def tail_norm(v):
    """_soc_tail_norm (soc_lorentz_kernels.jl:22-34): max-scaled two-pass norm."""
    scale = 0.0
    for i in range(1, len(v)):
        scale = max(scale, abs(v[i]))
    if scale == 0.0:
        return scale
    ss = 0.0
    for i in range(1, len(v)):
        t = v[i] / scale
        ss += t * t
    return scale * math.sqrt(ss)


def margin(v):
    """_soc_margin (soc_lorentz_kernels.jl:36)."""
    return v[0] - tail_norm(v)


def determinant(v):
    """_soc_determinant (soc.jl:114-120)."""
    d = v[0] * v[0]
    for i in range(1, len(v)):
        d -= v[i] * v[i]
    return d


def is_interior(v):
    """_soc_is_interior (soc.jl:122-123)."""
    return v[0] > 0.0 and margin(v) > 0.0


def sqrt_determinant(v):
    """_soc_sqrt_determinant (soc_lorentz_kernels.jl:42-62) -> (ok, sqrt det)."""
    scale = 0.0
    for x in v:
        if not math.isfinite(x):
            return False, 0.0
        scale = max(scale, abs(x))
    if scale == 0.0:
        return False, 0.0
    head = v[0]
    tail = tail_norm(v)
    mg = head - tail
    if not mg > 0.0:
        return False, 0.0
    residual = (mg / scale) * (head / scale + tail / scale)
    if not residual > 0.0:
        return False, 0.0
    r = scale * math.sqrt(residual)
    return (math.isfinite(r) and r > 0.0), r


def nt_scaling(s, z):
    """_soc_nt_scaling! (soc_lorentz_kernels.jl:72-184) -> (ok, w, lambda, eta, eta^2).

    On failure w and lambda are None (the reference commits no output)."""
    q = len(s)
    fail = (False, None, None, 0.0, 0.0)
    sok, ss = sqrt_determinant(s)
    zok, zs = sqrt_determinant(z)
    if not (sok and zok):
        return fail
    eta = math.sqrt(ss) / math.sqrt(zs)
    eta2 = eta * eta
    if not (math.isfinite(eta2) and eta2 > 0.0):
        return fail
    a = [0.0] * q
    b = [0.0] * q
    raw = [0.0] * q
    a[0] = s[0] / ss
    b[0] = z[0] / zs
    raw[0] = a[0] + b[0]
    for i in range(1, q):
        a[i] = s[i] / ss
        b[i] = z[i] / zs
        raw[i] = a[i] - b[i]
    if not all(math.isfinite(x) for x in a + b + raw):
        return fail
    ms = 1.0
    for i in range(q):
        ms = max(ms, abs(a[i]), abs(b[i]))
    ims = 1.0 / ms
    cross = (a[0] * ims) * (b[0] * ims)
    for i in range(1, q):
        cross += (a[i] * ims) * (b[i] * ims)
    radicand = ims * ims + cross
    if not (math.isfinite(radicand) and radicand > 0.0):
        return fail
    wscale = math.sqrt(2.0) * ms * math.sqrt(radicand)
    if not (math.isfinite(wscale) and wscale > 0.0):
        return fail
    nts = 1.0
    for i in range(1, q):
        nts = max(nts, abs(raw[i] / wscale))
    its = 1.0 / nts
    acc = its * its
    for i in range(1, q):
        t = (raw[i] / wscale) / nts
        acc += t * t
    nhead = nts * math.sqrt(acc)
    if not (math.isfinite(nhead) and nhead > 0.0):
        return fail
    gamma = wscale / 2.0
    ls = max(1.0, abs(a[0]), abs(b[0]), abs(gamma))
    for i in range(1, q):
        ls = max(ls, abs(a[i]), abs(b[i]))
    ils = 1.0 / ls
    aah = a[0] * ils
    bbh = b[0] * ils
    gg = gamma * ils
    den = aah + bbh + 2.0 * gg
    if not (math.isfinite(den) and den > 0.0):
        return fail
    geo = math.sqrt(ss) * math.sqrt(zs)
    lhead = geo * gamma
    if not math.isfinite(lhead):
        return fail
    lam = [0.0] * q
    for i in range(1, q):
        aa = a[i] * ils
        bb = b[i] * ils
        scaled = (gg + bbh) * aa + (gg + aah) * bb
        val = geo * (scaled / den) * ls
        if not math.isfinite(val):
            return fail
        lam[i] = val
    w = [0.0] * q
    w[0] = nhead
    lam[0] = lhead
    for i in range(1, q):
        w[i] = raw[i] / wscale
    return True, w, lam, eta, eta2


def apply_w(w, eta, src):
    """_soc_nt_apply_w! (soc_lorentz_kernels.jl:187-205)."""
    q = len(src)
    head = src[0]
    dt = 0.0
    for i in range(1, q):
        dt += w[i] * src[i]
    corr = head + dt / (1.0 + w[0])
    out = [0.0] * q
    out[0] = eta * (w[0] * head + dt)
    for i in range(1, q):
        out[i] = eta * (src[i] + corr * w[i])
    return out


def apply_winv(w, eta, src):
    """_soc_nt_apply_winv! (soc_lorentz_kernels.jl:208-226)."""
    q = len(src)
    head = src[0]
    dt = 0.0
    for i in range(1, q):
        dt += w[i] * src[i]
    corr = -head + dt / (1.0 + w[0])
    out = [0.0] * q
    out[0] = (w[0] * head - dt) / eta
    for i in range(1, q):
        out[i] = (src[i] + corr * w[i]) / eta
    return out


def apply_hs_inverse(w, eta2, src):
    """_soc_nt_apply_hs_inverse! (soc_lorentz_kernels.jl:229-248)."""
    q = len(src)
    head = src[0]
    ld = w[0] * head
    for i in range(1, q):
        ld -= w[i] * src[i]
    out = [0.0] * q
    out[0] = (2.0 * w[0] * ld - head) / eta2
    for i in range(1, q):
        out[i] = (-2.0 * w[i] * ld + src[i]) / eta2
    return out


def metric_matrix(w, eta2):
    """(W'W)^-1 of one cone as a q x q matrix: _soc_nt_apply_hs_inverse! on each unit vector, the
    dense branch of _native_soc_add_metric! (soc_native.jl:1180-1205)."""
    q = len(w)
    M = np.zeros((q, q))
    basis = [0.0] * q
    for j in range(q):
        basis[j] = 1.0
        M[:, j] = apply_hs_inverse(w, eta2, basis)
        basis[j] = 0.0
    return M


def add_metric(H, A, w, eta2):
    """H += A' (M A) for one cone with coefficient rows A (q x n): the dense single-cone assembly of
    _native_soc_add_metric! (soc_native.jl:1173-1217)."""
    M = metric_matrix(w, eta2)
    H += A.T @ (M @ A)
    return H


def jordan(left, right):
    """_soc_jordan! (soc.jl:94-112): x o y = (x'y, x0 y_tail + y0 x_tail)."""
    q = len(left)
    v = left[0] * right[0]
    for i in range(1, q):
        v += left[i] * right[i]
    out = [0.0] * q
    out[0] = v
    for i in range(1, q):
        out[i] = left[0] * right[i] + right[0] * left[i]
    return out


def jordan_solve(left, right):
    """_soc_jordan_solve! (soc_lorentz_kernels.jl:464-485): left o result = right.

    Returns None when ``left`` is not interior (the reference throws)."""
    if not is_interior(left):
        return None
    q = len(left)
    lh, rh = left[0], right[0]
    det = determinant(left)
    td = 0.0
    for i in range(1, q):
        td += left[i] * right[i]
    rhead = (lh * rh - td) / det
    ts = (td / lh - rh) / det
    out = [0.0] * q
    out[0] = rhead
    for i in range(1, q):
        out[i] = ts * left[i] + right[i] / lh
    return out


def fraction_to_boundary(s, ds):
    """_soc_fraction_to_boundary (soc.jl:148-214): largest t in [0,1] keeping s + t ds in the cone."""
    q = len(s)
    scale = 0.0
    for i in range(q):
        scale = max(scale, abs(s[i]), abs(ds[i]))
    if scale == 0.0:
        return 1.0
    s0 = s[0] / scale
    d0 = ds[0] / scale
    c0 = s0 * s0
    c1 = 2.0 * s0 * d0
    c2 = d0 * d0
    ff = s0 + d0
    fdet = ff * ff
    for i in range(1, q):
        si = s[i] / scale
        di = ds[i] / scale
        c0 -= si * si
        c1 -= 2.0 * si * di
        c2 -= di * di
        fv = si + di
        fdet -= fv * fv
    if s0 == 0.0 and d0 < 0.0:
        return 0.0
    if c0 == 0.0 and (c1 < 0.0 or (c1 == 0.0 and c2 < 0.0)):
        return 0.0
    if ff >= 0.0 and fdet >= 0.0:
        return 1.0
    root = 1.0
    if c2 == 0.0:
        if c1 < 0.0:
            root = min(root, -c0 / c1)
    else:
        disc = max(c1 * c1 - 4.0 * c2 * c0, 0.0)
        sq = math.sqrt(disc)
        qq = -(c1 + sq) / 2.0 if c1 >= 0.0 else -(c1 - sq) / 2.0
        if qq != 0.0:
            first = qq / c2
            second = c0 / qq
            if 0.0 < first <= root:
                root = first
            if 0.0 < second <= root:
                root = second
        else:
            rep = -c1 / (2.0 * c2)
            if 0.0 < rep <= root:
                root = rep
    if d0 < 0.0:
        root = min(root, -s0 / d0)
    return min(max(root, 0.0), 1.0)


def strict_step(state, direction, bound, allow_full_step=False):
    """_native_soc_strict_step (soc_native.jl:2705-2745) over a list of cones."""
    safety = 99.0 / 100.0
    cand = min(1.0, max(0.0, bound))
    if allow_full_step and cand == 1.0:
        if all(is_interior([x + d for x, d in zip(s, ds)]) for s, ds in zip(state, direction)):
            return 1.0
    step = safety * cand
    for _ in range(24):
        if all(is_interior([x + step * d for x, d in zip(s, ds)]) for s, ds in zip(state, direction)):
            return step
        step *= 1.0 / 2.0
    return 0.0


def lorentz_det3(x0, x1, x2):
    """_soc_lorentz_determinant (soc_lorentz_kernels.jl:251)."""
    return x0 * x0 - x1 * x1 - x2 * x2


def sym2_inverse_entries(x0, x1, x2):
    """_soc_sym2_inverse_entries (soc_lorentz_kernels.jl:263-266)."""
    d = lorentz_det3(x0, x1, x2)
    return (x0 - x1) / d, -x2 / d, (x0 + x1) / d


def hkm_metric(primal, dual):
    """_soc_fixed_trace_hkm_metric! (soc_lorentz_kernels.jl:353-370) -> (ok, 3-vector)."""
    x0, x1, x2 = primal
    z0, z1, z2 = dual
    tn = math.sqrt(x1 * x1 + x2 * x2)
    if not x0 > tn:
        return False, None
    d = lorentz_det3(x0, x1, x2)
    if not d > 0.0:
        return False, None
    return True, [(x0 * z0 - x1 * z1 + x2 * z2) / d, -(x1 * z2 + x2 * z1) / d, (x0 * z0 + x1 * z1 - x2 * z2) / d]


def _residual_r(x11, x12, x22, y11, y12, y22, target, aff, ap, ad):
    r11 = target - (x11 * y11 + x12 * y12)
    r12 = -(x11 * y12 + x12 * y22)
    r21 = -(x12 * y11 + x22 * y12)
    r22 = target - (x12 * y12 + x22 * y22)
    if aff:
        dx0, dx1, dx2 = ap
        dz0, dz1, dz2 = ad
        dx11, dx12, dx22 = dx0 + dx1, dx2, dx0 - dx1
        dy11, dy12, dy22 = (dz0 + dz1) / 2, dz2 / 2, (dz0 - dz1) / 2
        r11 -= dx11 * dy11 + dx12 * dy12
        r12 -= dx11 * dy12 + dx12 * dy22
        r21 -= dx12 * dy11 + dx22 * dy12
        r22 -= dx12 * dy12 + dx22 * dy22
    return r11, r12, r21, r22


def hkm_rhs_coordinates(primal, dual, presid, aff_p, aff_d, target, include_affine):
    """_soc_fixed_trace_hkm_rhs_coordinates (soc_lorentz_kernels.jl:372-418)."""
    x0, x1, x2 = primal
    z0, z1, z2 = dual
    p0, p1, p2 = presid
    x11, x12, x22 = x0 + x1, x2, x0 - x1
    y11, y12, y22 = (z0 + z1) / 2, z2 / 2, (z0 - z1) / 2
    p11, p12, p22 = p0 + p1, p2, p0 - p1
    r11, r12, r21, r22 = _residual_r(x11, x12, x22, y11, y12, y22, target, include_affine, aff_p, aff_d)
    w11 = p11 * y11 + p12 * y12 - r11
    w12 = p11 * y12 + p12 * y22 - r12
    w21 = p12 * y11 + p22 * y12 - r21
    w22 = p12 * y12 + p22 * y22 - r22
    i11, i12, i22 = sym2_inverse_entries(x0, x1, x2)
    q11 = i11 * w11 + i12 * w21
    q12 = i11 * w12 + i12 * w22
    q21 = i12 * w11 + i22 * w21
    q22 = i12 * w12 + i22 * w22
    return [q11 + q22, q11 - q22, q12 + q21]


def hkm_recovery(primal, dual, pdir, aff_p, aff_d, target, include_affine):
    """_soc_fixed_trace_hkm_recovery! (soc_lorentz_kernels.jl:420-461)."""
    x0, x1, x2 = primal
    z0, z1, z2 = dual
    dx0, dx1, dx2 = pdir
    x11, x12, x22 = x0 + x1, x2, x0 - x1
    y11, y12, y22 = (z0 + z1) / 2, z2 / 2, (z0 - z1) / 2
    dx11, dx12, dx22 = dx0 + dx1, dx2, dx0 - dx1
    r11, r12, r21, r22 = _residual_r(x11, x12, x22, y11, y12, y22, target, include_affine, aff_p, aff_d)
    w11 = r11 - (dx11 * y11 + dx12 * y12)
    w12 = r12 - (dx11 * y12 + dx12 * y22)
    w21 = r21 - (dx12 * y11 + dx22 * y12)
    w22 = r22 - (dx12 * y12 + dx22 * y22)
    i11, i12, i22 = sym2_inverse_entries(x0, x1, x2)
    q11 = i11 * w11 + i12 * w21
    q12 = i11 * w12 + i12 * w22
    q21 = i12 * w11 + i22 * w21
    q22 = i12 * w12 + i22 * w22
    return [q11 + q22, q11 - q22, q12 + q21]
