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
