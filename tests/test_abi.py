"""CPU checks of the C ABI: the HIP library loads without a GPU and exports
every entry point that include/*.h declares; the ctypes table matches."""
import ctypes
import glob
import os
import re

from sciwasm.sparse_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        for m in re.finditer(r"SABLE_API\s+[^;(]*?\b(sable_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "sable_spmv_csr" in names and "sable_ctx_create" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    missing = [n for n in sorted(declared_symbols()) if not hasattr(L, n)]
    assert not missing, missing


def test_ctypes_table_covers_header():
    assert declared_symbols() == set(_lib.SIGNATURES), \
        declared_symbols().symmetric_difference(set(_lib.SIGNATURES))


def test_version_and_error_string_without_gpu():
    L = _lib.load()
    assert b"gfx950" in L.sable_version()
    assert isinstance(L.sable_last_error(), bytes)


def test_null_arguments_are_rejected_without_touching_gpu():
    L = _lib.load()
    rc = L.sable_ctx_create(0, None, None)
    assert rc == -1
    assert b"NULL" in L.sable_last_error()
    rc = L.sable_spmv_csr(None, 1, 4, None, None, None, None, None, 1.0, 1)
    assert rc == -1
