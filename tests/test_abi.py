"""CPU checks of the C ABI: the built library loads and exports every symbol
that include/dss_gpu.h declares (no compute calls without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "distributed-string-sorting_amd", "dss_amd", "libdss_gpu.so")


def declared_symbols():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(dss_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_declared_symbols():
    assert os.path.exists(LIB), "libdss_gpu.so not built"
    lib = ctypes.CDLL(LIB)
    names = declared_symbols()
    assert len(names) > 10
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_declared_symbols():
    import dss_amd._lib as L
    assert set(L.symbols()) == declared_symbols()
