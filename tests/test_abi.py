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
