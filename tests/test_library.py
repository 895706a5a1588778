"""The C-ABI library builds, loads and exports every entry point declared in include/*.h
(no compute calls: these run without a GPU)."""

import ctypes
import os
import re

from helpers import ROOT, load_pkg


def declared_functions():
    names = []
    inc = os.path.join(ROOT, "include")
    for fn in sorted(os.listdir(inc)):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(skg_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_all_declared_symbols():
    pkg = load_pkg()
    assert os.path.exists(pkg.LIB_PATH), "libskirt_mi355x.so not built"
    lib = ctypes.CDLL(pkg.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert len(declared_functions()) >= 20


def test_python_binding_covers_header():
    pkg = load_pkg()
    assert set(pkg.EXPORTED) <= set(declared_functions())


def test_create_without_device_reports_error():
    """skg_create fails cleanly (status, message) when no HIP device is visible."""
    import pytest
    pkg = load_pkg()
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(pkg.EngineError):
        pkg.Engine(0)
