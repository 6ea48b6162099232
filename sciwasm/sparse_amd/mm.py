"""MatrixMarket ingest on the device.

Mirrors the reference's reader (seq/src/sparse32.js:1889-2172, 2498-2510;
FP64 seq/src/sparse64.js:1536-1760, 2051-2063):

* ``read_matrix_MM_files(text)`` -> ``MMInfo`` (the reference's mm_info:
  field, symmetry, nrows, ncols, nentries, nnz), with the entries parsed on
  the GPU and held there;
* ``create_COO_from_MM(mm_info, dtype)`` -> the full COO (symmetric
  entries expanded, explicit zeros dropped, sorted by (row, col));
* ``create_LCOO_from_MM(mm_info, dtype)`` -> the lower-triangular COO with
  a unit diagonal that the reference's triangular-solve tests build;
* ``mmread(filename, dtype)`` = read + create_COO_from_MM.

The text is copied to HBM once; line splitting, number parsing (rounded as
JavaScript's Number() rounds), expansion and sorting run on the device.
"""
import ctypes
import mmap
import os

import numpy as np

from . import _lib
from ._lib import call, c_vp, c_i64
from .core import Context, DeviceArray, sparse_COO_t, dtype_code

FIELDS = {0: "real", 1: "integer", 2: "complex", 3: "pattern", 4: "other"}


class MMInfo:
    """The reference's sparse_MM_info (seq/src/sparse32.js:14-25) plus the
    device-resident parsed entries.  Free with ``free()`` (or let it be
    collected)."""

    def __init__(self, handle, ctx):
        self.handle = handle
        self.ctx = ctx
        info = _lib.MMInfo()
        call("sable_mm_get_info", handle, ctypes.byref(info))
        self.field = FIELDS[info.field]
        self.symmetry = "symmetric" if info.symmetric else "general"
        self.nrows, self.ncols, self.nentries = info.nrows, info.ncols, info.nentries
        self.nnz = info.nnz
        self.lower_nnz = info.lower_nnz
        self.body_offset = info.body_offset

    def stats(self):
        """(lines holding an entry or a blank, values converted exactly)."""
        lines, exact = c_i64(), c_i64()
        call("sable_mm_stats", self.handle, ctypes.byref(lines), ctypes.byref(exact))
        return lines.value, exact.value

    def free(self):
        if self.handle:
            _lib.load().sable_mm_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _as_bytes(text):
    if isinstance(text, str):
        return text.encode()
    return text


def read_matrix_MM_files(text, ctx=None):
    """Parse the whole text of a coordinate MatrixMarket file.

    text: str / bytes / bytearray / memoryview / mmap on the host, or a
    DeviceArray of uint8 already in HBM."""
    ctx = ctx if ctx is not None else Context.default()
    h = c_vp()
    if isinstance(text, DeviceArray):
        call("sable_mm_parse", ctx.handle, text.ptr, text.nbytes, 1, ctypes.byref(h))
    else:
        buf = np.frombuffer(_as_bytes(text), dtype=np.uint8)
        if buf.size == 0:
            raise _lib.SableError(-3, "sable_mm_parse", "empty text")
        call("sable_mm_parse", ctx.handle, buf.ctypes.data, buf.size, 0, ctypes.byref(h))
    return MMInfo(h, ctx)


def _coo(mm_info, dtype, lower):
    ctx = mm_info.ctx
    n = mm_info.lower_nnz if lower else mm_info.nnz
    row = DeviceArray(n, np.int32, ctx)
    col = DeviceArray(n, np.int32, ctx)
    val = DeviceArray(n, dtype, ctx)
    call("sable_mm_coo", ctx.handle, mm_info.handle, dtype_code(dtype), int(lower), row.ptr, col.ptr, val.ptr)
    return sparse_COO_t(row, col, val, mm_info.nrows, n)


def create_COO_from_MM(mm_info, dtype=np.float32):
    """seq/src/sparse32.js:2100-2172 (FP64: seq/src/sparse64.js)."""
    return _coo(mm_info, np.dtype(dtype), False)


def create_LCOO_from_MM(mm_info, dtype=np.float32):
    """seq/src/sparse32.js:1983-2097, par/src/spts32.js:14."""
    return _coo(mm_info, np.dtype(dtype), True)


def mmread(filename, dtype=np.float32, ctx=None):
    """seq/src/sparse32.js:2498-2510: file -> COO on the device.  The file
    is memory-mapped and copied to HBM once."""
    with open(filename, "rb") as f:
        if os.fstat(f.fileno()).st_size == 0:
            raise _lib.SableError(-3, "mmread", "empty file")
        with mmap.mmap(f.fileno(), 0, access=mmap.ACCESS_READ) as mm:
            info = read_matrix_MM_files(mm, ctx)
    try:
        return create_COO_from_MM(info, dtype)
    finally:
        info.free()
