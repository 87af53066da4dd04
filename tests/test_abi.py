"""The C-ABI library loads on a machine without a GPU and exports every
entry point that include/dns_rocm.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

from distributednavierstokes_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(dns_\w+)\s*\(", text, re.M))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "dns_fv_create" in names and "dns_fv_step" in names and "dns_last_error" in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIBPATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_symbols()) <= set(_lib.exported_symbols())


def test_version_and_error_without_gpu():
    L = _lib.lib()
    assert b"gfx950" in L.dns_version()
    assert L.dns_fv_create(None, None) == -1
    assert b"null" in L.dns_last_error()
