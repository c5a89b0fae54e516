"""The C-ABI library loads and exports every symbol include/pfhip.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "pfhip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|int32_t|const char\*)\s+(pf_\w+)\s*\(", text, re.M)))


def test_header_declares_the_entry_points():
    syms = declared_symbols()
    for s in ("pf_create", "pf_destroy", "pf_advance", "pf_read_field", "pf_dft_read",
              "pf_last_error", "pf_device_info"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from photonfdtd_amd import _native
    lib = _native.load_library()
    for s in declared_symbols():
        assert hasattr(lib, s), f"{s} not exported"
        assert s in _native.SIGNATURES, f"{s} has no ctypes signature"
    assert lib.pf_abi_version() == _native.ABI_VERSION


def test_invalid_descriptor_is_rejected_without_touching_the_gpu():
    from photonfdtd_amd import _native as N
    lib = N.load_library()
    d = N.PfDesc()
    d.dtype = 7
    h = ctypes.c_void_p()
    rc = lib.pf_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == -1
    assert b"shape" in lib.pf_last_error() or b"dtype" in lib.pf_last_error()
    assert lib.pf_advance(None, 0, 1) == -1
    assert lib.pf_last_error() == b"null handle"


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    from photonfdtd_amd import _native
    monkeypatch.setattr(_native, "_LIB", None)
    with pytest.raises(RuntimeError, match="missing"):
        _native.load_library(tmp_path / "nope.so")
