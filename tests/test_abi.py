"""The C-ABI libraries load and export every symbol include/pt_hip.h and
include/pt_h5.h declare (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from particle_transport_amd import _lib, h5

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header="pt_hip.h"):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert "pt_ens_advance" in syms and "pt_dmumu_accumulate" in syms
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_declarations_and_abi_version():
    L = _lib.load()
    assert set(declared_symbols()) <= set(_lib.EXPORTS) | {"pt_abi_version"}
    assert L.pt_abi_version() == 1


def test_errors_are_reported_not_thrown():
    L = _lib.load()
    # a null out-pointer is rejected with a message before any device call
    rc = L.pt_ctx_create(0, None)
    assert rc == _lib.PT_ERR_INVALID
    assert b"null" in L.pt_last_error()


def test_h5_library_exports_every_declared_symbol():
    syms = declared_symbols("pt_h5.h")
    assert "pt_h5_write_box" in syms and "pt_h5_read_box" in syms
    lib = ctypes.CDLL(h5.LIB_PATH)
    assert not [s for s in syms if not hasattr(lib, s)]
    assert set(syms) == set(h5.EXPORTS)
