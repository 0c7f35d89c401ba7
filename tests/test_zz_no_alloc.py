"""sdpx_soc_fraction_to_boundary_noalloc against sdpx_soc_fraction_to_boundary, through the C-ABI.

The entry point that allocates nothing runs the same two kernels as the
handle's entry point (per-cone bound, then the fixed-order minimum) on
caller-owned device buffers, with the per-cone bounds in a static device array
when the caller passes none.  Both must give the same bits.  Two calls in a
row, the second on fewer cones, check that the static array's leftover
contents never leak into a result.
"""

import ctypes as C

import numpy as np
import pytest

from sdpx_mi355x import _lib
from sdpx_mi355x.soc import LorentzCones, random_interior


def _cones(rng, n):
    dims = rng.choice([1, 2, 3, 4, 5, 7, 8, 13], size=n).astype(np.int32)
    s = random_interior(rng, dims, 1.5)
    ds = rng.standard_normal(int(dims.sum())) * 2.0
    return dims, s, ds


def _handle_result(dims, s, ds):
    with LorentzCones(dims) as c:
        c.set(0, s)
        c.set(1, ds)
        return c.fraction_to_boundary(0, 1, per_cone=True)


class _Hip:
    """Device buffers from the HIP runtime the library itself is linked against (libamdhip64.so.7)."""

    H2D, D2H = 1, 2

    def __init__(self):
        _lib.load()
        self.rt = C.CDLL("libamdhip64.so.7")
        self.ptrs = []

    def _ok(self, rc, what):
        assert rc == 0, f"{what} failed: {rc}"

    def upload(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self._ok(self.rt.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 8))), "hipMalloc")
        self.ptrs.append(p.value)
        self._ok(self.rt.hipMemcpy(C.c_void_p(p.value), a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes),
                                   self.H2D), "hipMemcpy")
        return p.value

    def download(self, ptr: int, n: int) -> np.ndarray:
        out = np.empty(n)
        self._ok(self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), C.c_size_t(out.nbytes),
                                   self.D2H), "hipMemcpy")
        return out

    def stream(self) -> int:
        st = C.c_void_p()
        self._ok(self.rt.hipStreamCreate(C.byref(st)), "hipStreamCreate")
        self.streams = getattr(self, "streams", []) + [st.value]
        return st.value

    def sync(self, stream: int):
        self._ok(self.rt.hipStreamSynchronize(C.c_void_p(stream)), "hipStreamSynchronize")

    def close(self):
        for p in self.ptrs:
            self.rt.hipFree(C.c_void_p(p))
        for st in getattr(self, "streams", []):
            self.rt.hipStreamDestroy(C.c_void_p(st))
        self.ptrs = []


def _noalloc(lib, hip, stream, dims, s, ds, per_cone: bool):
    offsets = hip.upload(np.concatenate([[0], np.cumsum(dims, dtype=np.int64)]).astype(np.int64))
    s_d, ds_d = hip.upload(s), hip.upload(ds)
    out = hip.upload(np.array([-1.0]))
    pc = hip.upload(np.full(dims.size, -1.0)) if per_cone else None
    print(f"noalloc call: {dims.size} cones, per_cone={per_cone}", flush=True)
    rc = lib.sdpx_soc_fraction_to_boundary_noalloc(0, int(dims.size), offsets, s_d, ds_d, out, pc, stream)
    assert rc == 0, lib.sdpx_last_error()
    hip.sync(stream)
    return float(hip.download(out, 1)[0]), (hip.download(pc, dims.size) if per_cone else None)


@pytest.mark.gpu
def test_noalloc_fraction_to_boundary_matches_handle(gpu_lib):
    lib = _lib.load()
    hip = _Hip()
    try:
        stream = hip.stream()
        rng = np.random.default_rng(11)
        dims, s, ds = _cones(rng, 300_000)
        ref_min, ref_pc = _handle_result(dims, s, ds)
        assert 0.0 < ref_min < 1.0
        got_min, _ = _noalloc(lib, hip, stream, dims, s, ds, per_cone=False)
        assert got_min == ref_min
        # second call, smaller input: the static array still holds the first call's 300 000 bounds
        n2 = 70_001
        dims2 = dims[:n2]
        t2 = int(dims2.sum())
        ds2 = ds[:t2] * 0.25  # other bounds than the first call's prefix
        ref2_min, _ = _handle_result(dims2, s[:t2], ds2)
        got2_min, _ = _noalloc(lib, hip, stream, dims2, s[:t2], ds2, per_cone=False)
        assert got2_min == ref2_min
        assert ref2_min != ref_min
        # the default stream, and a caller-provided per-cone array: same bits per cone
        got3_min, got3_pc = _noalloc(lib, hip, 0, dims, s, ds, per_cone=True)
        assert got3_min == ref_min
        assert np.array_equal(got3_pc, ref_pc)
    finally:
        hip.close()


def test_noalloc_rejects_too_many_cones():
    """Host-side argument check; nothing is launched."""
    lib = _lib.load()
    cap = 1 << 22
    dummy = C.c_double(0.0)
    rc = lib.sdpx_soc_fraction_to_boundary_noalloc(0, cap + 1, C.addressof(dummy), C.addressof(dummy),
                                                   C.addressof(dummy), C.addressof(dummy), None, None)
    assert rc == -1
    assert b"scratch capacity" in lib.sdpx_last_error()
    rc = lib.sdpx_soc_fraction_to_boundary_noalloc(0, 0, None, None, None, None, None, None)
    assert rc == -1
