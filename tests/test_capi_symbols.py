"""The C-ABI library loads and exports every symbol include/axs3d.h
declares (no device needed; nothing is computed)."""
import ctypes
import os
import re

from axisem3d_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "axs3d.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(axs3d_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_table():
    assert _declared() == sorted(capi.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_errors_without_context():
    lib = capi.load()
    assert lib.axs3d_last_error(None) == b"axs3d || Null context."
    assert lib.axs3d_finalize(None) != 0
