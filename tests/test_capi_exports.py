"""The C-ABI library loads and exports every symbol its headers declare (CPU only)."""
import ctypes
import os
import re

import feelpp_amd as fa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(feel_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = ctypes.CDLL(fa.LIBPATH)
    names = declared("feelpp_hip.h") + declared("feelpp_host.h")
    assert len(names) > 50
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_bindings_cover_the_headers():
    names = set(declared("feelpp_hip.h") + declared("feelpp_host.h"))
    assert names == set(fa._SIGS), names ^ set(fa._SIGS)


def test_last_error_is_empty_string_initially():
    assert fa.lib().feel_hip_last_error() is not None
