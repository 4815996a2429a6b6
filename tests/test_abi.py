"""The C ABI library loads on a machine without a GPU and exports every entry
point that include/racon_hip.h declares; the ctypes binding covers them all.
No compute call is made here."""
import ctypes
import os
import re
import subprocess

from racon_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "racon_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rh_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "rh_poa_run" in syms and "rh_align_run" in syms
    assert len(syms) >= 15


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (rh_[a-z0-9_]+)\b", out))
    assert set(declared_symbols()) <= exported


def test_binding_covers_header():
    assert set(declared_symbols()) == set(_lib._SIGNATURES)


def test_load_library_and_version():
    lib = _lib.load_library()
    assert lib.rh_version() >= 100
    assert lib.rh_last_error() is not None


def test_missing_library_fails_loudly(tmp_path):
    saved = _lib._lib
    try:
        _lib._lib = None
        try:
            _lib.load_library(str(tmp_path / "missing.so"))
        except RuntimeError as e:
            assert "not found" in str(e)
        else:
            raise AssertionError("load_library accepted a missing library")
    finally:
        _lib._lib = saved
