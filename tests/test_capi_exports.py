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


def test_null_simulation_is_refused_without_a_gpu():
    """Every entry point that takes a simulation handle refuses a null one with
    LBM3_E_ARG and a message, before any device call (so no GPU is needed)."""
    import re

    L = lib.load()
    with open(lib.HEADER) as fh:
        header = fh.read()
    names = re.findall(r"^int (lbm3_\w+)\(lbm3_sim \*s[,)]", header, re.M)
    names = [n for n in names if n not in ("lbm3_sim_destroy", "lbm3_level_count")]
    assert len(names) >= 25
    for n in names:
        fn = getattr(L, n)
        args = []
        for t in fn.argtypes:
            if t is ctypes.c_char_p:
                args.append(b"unused")
            elif t in (ctypes.c_int, ctypes.c_longlong, ctypes.c_uint, ctypes.c_ulonglong):
                args.append(1)
            elif t in (ctypes.c_float, ctypes.c_double):
                args.append(0.0)
            elif isinstance(t, type) and issubclass(t, ctypes.Structure):
                args.append(t())
            else:
                args.append(None)
        assert fn(*args) == -1, n
        assert L.lbm3_last_error(), n
    assert L.lbm3_sim_destroy(None) in (0, -1)
