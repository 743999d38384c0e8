"""libmisohip.so loads on the CPU and exports every entry point that
include/misohip.h declares (no compute call is made without a GPU)."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            names += re.findall(r"^\s*(?:int|const char \*)\s*(misohip_\w+)\s*\(", text, re.M)
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "misohip_step" in names and "misohip_rk_stage" in names
    assert len(names) == len(set(names))


def test_library_exports_every_declared_symbol():
    from miso_amd import hip
    lib = C.CDLL(hip.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from miso_amd import hip
    assert set(declared_symbols()) == set(hip.SIGNATURES)


def test_strerror_without_gpu():
    from miso_amd import hip
    L = hip.lib()
    assert L.misohip_strerror(0) == b"success"
    assert b"argument" in L.misohip_strerror(-2)
    # argument validation happens before any device work
    assert L.misohip_create(0, None, None) == -2
