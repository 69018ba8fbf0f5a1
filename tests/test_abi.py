"""The C ABI: libbve.so loads and exports every entry point include/bve.h
declares, with the signatures the ctypes layer uses (no GPU needed)."""
import ctypes
import os
import re

from conftest import ROOT
from thesiscode_amd import _lib


def header_functions():
    text = open(os.path.join(ROOT, "include", "bve.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bve_\w+)\s*\(", text)))


def test_header_declares_the_python_signatures():
    assert header_functions() == sorted(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    for name in header_functions():
        assert hasattr(L, name), name


def test_create_without_device_fails_cleanly():
    """On a host with no HIP device bve_create returns an error code and a
    message instead of crashing (on a GPU host it succeeds)."""
    L = _lib.load()
    h = ctypes.c_void_p()
    rc = L.bve_create(ctypes.byref(h), 0)
    if rc == 0:
        assert L.bve_destroy(h) == 0
    else:
        assert rc == 3 and b"device" in L.bve_last_error(None)
        assert not h.value
