"""The C-ABI library loads (no GPU needed) and exports every entry point that
include/lbm3mi.h declares."""
import ctypes

from lbm3_amd import lib


def test_library_exports_every_declared_symbol():
    so = ctypes.CDLL(lib.LIB_PATH)
    declared = lib.exported_symbols_in_header()
    assert len(declared) >= 20
    missing = [s for s in declared if not hasattr(so, s)]
    assert not missing, missing


def test_binding_covers_header():
    L = lib.load()
    for s in lib.exported_symbols_in_header():
        assert getattr(L, s).restype is not None or s == "lbm3_last_error"
