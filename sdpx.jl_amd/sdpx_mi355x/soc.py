"""Host binding of the Lorentz-cone kernels (sdpx_soc_* in include/sdpx_mi355x.h).

Mirrors the native SOC workspace's per-cone kernel calls
(src/soc_native.jl over src/soc_lorentz_kernels.jl and src/soc.jl): the cones'
coordinate vectors live on the device in numbered slots, every method is one
launch over all cones.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dptr

SLOTS = 8
APPLY_W, APPLY_WINV, APPLY_HS_INVERSE = 0, 1, 2


class LorentzCones:
    """Device-resident cones of dimensions ``dims`` (concatenated, head first)."""

    def __init__(self, dims, device: int = 0):
        lib = _lib.load()
        self.dims = np.ascontiguousarray(dims, dtype=np.int32)
        self.ncones = int(self.dims.size)
        self.total = int(self.dims.sum())
        self.ptr = np.concatenate([[0], np.cumsum(self.dims)]).astype(np.int64)
        h = C.c_void_p()
        check(lib.sdpx_soc_create(device, self.ncones, self.dims.ctypes.data_as(_lib._ip), C.byref(h)),
              "sdpx_soc_create")
        self._h = h
        self._lib = lib

    def close(self):
        if self._h:
            self._lib.sdpx_soc_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- data movement ---------------------------------------------------------------
    def set(self, slot: int, v) -> None:
        a = np.ascontiguousarray(v, dtype=np.float64)
        if a.size != self.total:
            raise ValueError(f"slot vector needs {self.total} values, got {a.size}")
        check(self._lib.sdpx_soc_set_vector(self._h, slot, dptr(a)), "sdpx_soc_set_vector")

    def get(self, slot: int) -> np.ndarray:
        a = np.empty(self.total)
        check(self._lib.sdpx_soc_get_vector(self._h, slot, dptr(a)), "sdpx_soc_get_vector")
        return a

    def copy(self, src: int, dst: int) -> None:
        check(self._lib.sdpx_soc_copy_vector(self._h, src, dst), "sdpx_soc_copy_vector")

    def scaling(self):
        eta, eta2 = np.empty(self.ncones), np.empty(self.ncones)
        ok = np.empty(self.ncones, dtype=np.int32)
        check(self._lib.sdpx_soc_get_scaling(self._h, dptr(eta), dptr(eta2), ok.ctypes.data_as(_lib._ip)),
              "sdpx_soc_get_scaling")
        return eta, eta2, ok.astype(bool)

    # --- kernels -------------------------------------------------------------------------
    def nt_scaling(self, s: int, z: int, w: int, lam: int) -> int:
        nf = C.c_int(0)
        check(self._lib.sdpx_soc_nt_scaling(self._h, s, z, w, lam, C.byref(nf)), "sdpx_soc_nt_scaling")
        return nf.value

    def apply(self, op: int, w: int, src: int, dst: int) -> None:
        check(self._lib.sdpx_soc_apply(self._h, op, w, src, dst), "sdpx_soc_apply")

    def jordan(self, left: int, right: int, dst: int) -> None:
        check(self._lib.sdpx_soc_jordan(self._h, left, right, dst), "sdpx_soc_jordan")

    def jordan_solve(self, left: int, right: int, dst: int) -> int:
        nb = C.c_int(0)
        check(self._lib.sdpx_soc_jordan_solve(self._h, left, right, dst, C.byref(nb)), "sdpx_soc_jordan_solve")
        return nb.value

    def inverse(self, src: int, dst: int) -> int:
        nb = C.c_int(0)
        check(self._lib.sdpx_soc_inverse(self._h, src, dst, C.byref(nb)), "sdpx_soc_inverse")
        return nb.value

    def fraction_to_boundary(self, s: int, ds: int, per_cone: bool = False):
        mn = C.c_double(0.0)
        pc = np.empty(self.ncones) if per_cone else None
        check(self._lib.sdpx_soc_fraction_to_boundary(self._h, s, ds, C.byref(mn), dptr(pc)),
              "sdpx_soc_fraction_to_boundary")
        return (mn.value, pc) if per_cone else mn.value

    def set_constraints(self, A) -> None:
        """A: (sum q) x n coefficient rows of every cone, stacked in cone order (uploaded once)."""
        A = np.asfortranarray(A, dtype=np.float64)
        if A.ndim != 2 or A.shape[0] != self.total:
            raise ValueError(f"A must have {self.total} rows (one per cone coordinate)")
        self.nvars = int(A.shape[1])
        check(self._lib.sdpx_soc_set_constraints(self._h, self.nvars, dptr(A)), "sdpx_soc_set_constraints")

    def metric(self, w: int, H0=None, download: bool = True):
        """H0 + sum_c A_c' (W_c'W_c)^-1 A_c -> (H or None, failed cones)."""
        n = self.nvars
        h0 = None if H0 is None else np.asfortranarray(H0, dtype=np.float64)
        if h0 is not None and h0.shape != (n, n):
            raise ValueError(f"H0 must be {n} x {n}")
        H = np.empty((n, n), order="F") if download else None
        nf = C.c_int(0)
        rc = self._lib.sdpx_soc_metric(self._h, w, dptr(h0), dptr(H), C.byref(nf))
        if rc < 0:
            check(rc, "sdpx_soc_metric")
        return H, nf.value

    def margins(self, slot: int):
        mg = np.empty(self.ncones)
        ni = C.c_int(0)
        check(self._lib.sdpx_soc_margins(self._h, slot, dptr(mg), C.byref(ni)), "sdpx_soc_margins")
        return mg, ni.value

    def strict_step(self, state: int, direction: int, bound: float, allow_full_step: bool = False) -> float:
        st = C.c_double(0.0)
        check(self._lib.sdpx_soc_strict_step(self._h, state, direction, bound, int(allow_full_step), C.byref(st)),
              "sdpx_soc_strict_step")
        return st.value

    def hkm_metric(self, primal: int, dual: int, dst: int) -> int:
        nf = C.c_int(0)
        check(self._lib.sdpx_soc_hkm_metric(self._h, primal, dual, dst, C.byref(nf)), "sdpx_soc_hkm_metric")
        return nf.value

    def hkm_rhs_coordinates(self, primal, dual, presid, aff_p, aff_d, target, include_affine, dst) -> None:
        check(self._lib.sdpx_soc_hkm_contract(self._h, 0, primal, dual, presid, aff_p, aff_d, target,
                                              int(include_affine), dst), "sdpx_soc_hkm_contract")

    def hkm_recovery(self, primal, dual, pdir, aff_p, aff_d, target, include_affine, dst) -> None:
        check(self._lib.sdpx_soc_hkm_contract(self._h, 1, primal, dual, pdir, aff_p, aff_d, target,
                                              int(include_affine), dst), "sdpx_soc_hkm_contract")

    def set_profiling(self, enabled: bool) -> None:
        check(self._lib.sdpx_soc_set_profiling(self._h, int(enabled)), "sdpx_soc_set_profiling")

    def last_kernel_ms(self) -> float:
        ms = C.c_double(0.0)
        check(self._lib.sdpx_soc_last_kernel_ms(self._h, C.byref(ms)), "sdpx_soc_last_kernel_ms")
        return ms.value

    def synchronize(self) -> None:
        check(self._lib.sdpx_soc_synchronize(self._h), "sdpx_soc_synchronize")


def random_interior(rng, dims, spread: float = 1.0) -> np.ndarray:
    """Strictly interior cone points: tail ~ N(0, 1), head = ||tail|| (1 + u) + 1e-3, u ~ U(0.01, spread)."""
    dims = np.asarray(dims, dtype=np.int64)
    total = int(dims.sum())
    v = rng.standard_normal(total)
    starts = np.concatenate([[0], np.cumsum(dims)[:-1]])
    sq = v * v
    sq[starts] = 0.0
    norms = np.sqrt(np.add.reduceat(sq, starts))
    v[starts] = norms * (1.0 + rng.uniform(0.01, spread, dims.size)) + 1e-3
    return v
