"""The C-ABI library loads without a GPU and exports every entry point the header declares."""

import ctypes
import pathlib
import re

import pytest

from aquaflux_amd import _lib

HEADER = pathlib.Path(__file__).resolve().parents[1] / "include" / "aquaflux_amd.h"


def header_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(afx_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_functions():
    names = header_functions()
    assert "afx_flow_residual" in names and "afx_mesh_create" in names
    assert len(names) >= 10


def test_library_exports_every_declared_symbol():
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(set(_lib.exported_symbols())) == header_functions()


def test_error_path_without_gpu():
    """Invalid arguments are rejected with status 2 and a message, before any device call."""
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    L = _lib.lib()
    h = ctypes.c_void_p()
    rc = L.afx_mesh_create(4, 1, 1, 1, None, None, None, None, ctypes.byref(h))
    assert rc == 2
    assert b"dim" in L.afx_last_error()
