"""The C-ABI library loads on the host and exports every symbol declared in include/*.h."""

import ctypes
import re
from pathlib import Path

import pytest

from sdpx_mi355x import _lib

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(sdpx_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "sdpx_dense_create" in names and "sdpx_dense_schur_build" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_symbols()) <= set(_lib.SIGNATURES)


def test_host_calls_without_gpu():
    lib = _lib.load()
    n = ctypes.c_int(-1)
    assert lib.sdpx_device_count(ctypes.byref(n)) == 0
    assert n.value >= 0
    assert lib.sdpx_version() >= 100
    out = ctypes.c_void_p()
    # invalid dimensions are rejected before any device call
    karr = (ctypes.c_int * 1)(0)
    assert lib.sdpx_dense_create(0, 1, karr, 5, ctypes.byref(out)) == -1
    assert b"invalid" in lib.sdpx_last_error() or b"positive" in lib.sdpx_last_error()


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_LIB", None)
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "absent.so")
    with pytest.raises(_lib.SDPXError):
        _lib.load()


def test_batch_inputs_checked_before_the_abi():
    """Mis-shaped member arrays raise ValueError on the host, before any copy reads past them."""
    import numpy as np

    from sdpx_mi355x.batch import BatchSolver

    bs = BatchSolver.__new__(BatchSolver)  # no device handle: every check below runs first
    bs.k, bs.m, bs.B = np.array([3, 4], np.int32), np.array([5, 6], np.int32), 2
    good_Av = [np.zeros((5, 9)), np.zeros((6, 16))]
    good_C = [np.eye(3), np.eye(4)]
    good_c = [np.zeros(5), np.zeros(6)]
    with pytest.raises(ValueError, match="member 1 has 80 values"):
        bs.set_dense([good_Av[0], np.zeros((5, 16))], good_C, good_c)
    with pytest.raises(ValueError, match="1 members given"):
        bs.set_dense(good_Av[:1], good_C, good_c)
    with pytest.raises(ValueError, match="set_start X"):
        bs.set_start(good_c, [np.eye(3), np.eye(3)], good_C)
    with pytest.raises(ValueError, match="adjoint h"):
        bs.adjoint(good_C, [np.zeros(5), np.zeros(5)])
    rows = [np.array([0], np.int32)] * 11
    with pytest.raises(ValueError, match="different lengths"):
        bs.set_coo(rows, rows, [np.array([1.0, 2.0])] * 11, good_C, good_c)
    with pytest.raises(ValueError, match="one entry list per constraint"):
        bs.set_coo(rows[:10], rows[:10], [np.array([1.0])] * 10, good_C, good_c)


def _block_coo(mats, k):
    """SparseBlockCOO of build_block_coo (types/constraints.jl:79-109): CSC order, 1-based, both triangles."""
    import numpy as np

    ptr, lin, row, col, val = [1], [], [], [], []
    for A in mats:
        for c in range(k):
            for r in range(k):
                if A[r, c] != 0.0:
                    row.append(r + 1)
                    col.append(c + 1)
                    lin.append(c * k + r + 1)
                    val.append(A[r, c])
        ptr.append(len(row) + 1)
    as32 = lambda v: np.array(v, dtype=np.int32)  # noqa: E731
    return as32(ptr), as32(lin), as32(row), as32(col), np.array(val, dtype=np.float64)


def _convert(k, m, ptr, lin, row, col, val):
    import numpy as np

    lib = _lib.load()
    n = max(int(ptr[-1]) - 1, 1)
    optr = np.zeros(m + 1, dtype=np.int64)
    orow, ocol = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)
    oval, nnz = np.zeros(n), ctypes.c_longlong(0)
    ip = ctypes.POINTER(ctypes.c_int)
    rc = lib.sdpx_block_coo_convert(k, m, ptr.ctypes.data_as(ip), lin.ctypes.data_as(ip), row.ctypes.data_as(ip),
                                    col.ctypes.data_as(ip), _lib.dptr(val),
                                    optr.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), orow.ctypes.data_as(ip),
                                    ocol.ctypes.data_as(ip), _lib.dptr(oval), ctypes.byref(nnz))
    return rc, optr, orow[: nnz.value], ocol[: nnz.value], oval[: nnz.value]


def test_sparse_block_coo_as_stored_converts_to_upper_coo():
    """The reference's SparseBlockCOO (1-based Int32, both triangles, lin) becomes the ABI's 0-based
    upper-triangle COO; the dense matrices both describe are equal."""
    import numpy as np

    k = 4
    A1 = np.zeros((k, k))
    A1[0, 0], A1[1, 3], A1[3, 1], A1[2, 2] = 2.0, -1.5, -1.5, 0.25
    A2 = np.zeros((k, k))
    A2[0, 2] = A2[2, 0] = 3.0
    A3 = np.zeros((k, k))  # an empty position
    ptr, lin, row, col, val = _block_coo([A1, A2, A3], k)
    rc, optr, orow, ocol, oval = _convert(k, 3, ptr, lin, row, col, val)
    assert rc == 0
    assert list(optr) == [0, 3, 4, 4]
    assert np.all(orow <= ocol) and np.all(orow >= 0) and np.all(ocol < k)
    for i, A in enumerate((A1, A2, A3)):
        D = np.zeros((k, k))
        for e in range(optr[i], optr[i + 1]):
            D[orow[e], ocol[e]] += oval[e]
            if orow[e] != ocol[e]:
                D[ocol[e], orow[e]] += oval[e]
        assert np.array_equal(D, A)


def test_sparse_block_coo_rejects_malformed_input():
    import numpy as np

    k = 3
    A = np.zeros((k, k))
    A[0, 1] = A[1, 0] = 1.0
    good = _block_coo([A], k)
    bad_lin = [a.copy() for a in good]
    bad_lin[1][0] += 1
    asym = [a.copy() for a in good]
    asym[4][1] = 2.0
    one_sided = _block_coo([np.triu(A)], k)
    zero_based = [a.copy() for a in good]
    zero_based[0] -= 1
    assert _convert(k, 1, *good)[0] == 0
    for case in (bad_lin, asym, one_sided, zero_based):
        assert _convert(k, 1, *case)[0] == -1
    lib = _lib.load()
    assert lib.sdpx_last_error()
