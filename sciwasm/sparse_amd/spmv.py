"""SpMV dispatch, mirroring spmv() of seq/src/sparse32.js:1232-1283.

``spmv(A, x, y, inner_max)`` computes y += A.x, inner_max times, on the GPU
for every format; ``spmv(A, x, y, beta=0)`` is the plain product.
"""
import ctypes
import sys

import numpy as np

from . import _lib
from ._lib import call, c_vp, c_i64, c_i32
from .core import (sparse_COO_t, sparse_CSR_t, sparse_DIA_t, sparse_ELL_t, dtype_code)


class CsrPlan:
    """Device row-block schedule for a CSR matrix (sable_csr_plan_*)."""

    def __init__(self, A, tune=None):
        self.ctx = A.row.ctx
        self.A = A
        h = c_vp()
        t = None
        if tune is not None:
            t = ctypes.byref(_lib.CsrTune(int(tune.get("block_weight", 0)), int(tune.get("row_weight", 0)),
                                          int(tune.get("vector_lanes", 0)), int(tune.get("vector_rows", 0))))
        call("sable_csr_plan_create", self.ctx.handle, A.N, A.row.ptr, t, ctypes.byref(h))
        self.handle = h

    def info(self):
        nb, nnz, cap = c_i64(), c_i64(), c_i32()
        call("sable_csr_plan_info", self.handle, ctypes.byref(nb), ctypes.byref(nnz), ctypes.byref(cap))
        return {"nblocks": nb.value, "nnz": nnz.value, "lds_cap": cap.value}

    def run(self, x, y, beta=1.0, inner_max=1):
        A = self.A
        call("sable_spmv_csr_plan", self.ctx.handle, self.handle, dtype_code(A.dtype), A.col.ptr, A.val.ptr,
             x.ptr, y.ptr, float(beta), int(inner_max))

    def __del__(self):
        try:
            if self.handle:
                _lib.load().sable_csr_plan_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class CsrTiles:
    """Column-sorted tile plan (sable_csr_tiles_*): the matrix copied into
    (column panel, row tile) groups sorted by column, rows summed in LDS,
    one launch per panel."""

    def __init__(self, A, tile_rows=0, panel_cols=0):
        self.ctx = A.row.ctx
        self.A = A
        h = c_vp()
        call("sable_csr_tiles_create", self.ctx.handle, dtype_code(A.dtype), A.N, A.row.ptr, A.col.ptr, A.val.ptr,
             int(tile_rows), int(panel_cols), ctypes.byref(h))
        self.handle = h

    def info(self):
        nt, npan, b = c_i32(), c_i32(), c_i64()
        call("sable_csr_tiles_info", self.handle, ctypes.byref(nt), ctypes.byref(npan), ctypes.byref(b))
        return {"row_tiles": nt.value, "npanels": npan.value, "bytes": b.value}

    def run(self, x, y, beta=1.0, inner_max=1):
        call("sable_spmv_csr_tiles", self.ctx.handle, self.handle, x.ptr, y.ptr, float(beta), int(inner_max))

    @classmethod
    def from_coo(cls, A, tile_rows=0, panel_cols=0):
        """A tile plan of a COO.  Row-sorted: its row heads (from a CooPlan)
        with its col and val arrays form the CSR the plan copies; otherwise
        the COO is first sorted on the device (coo_csr)."""
        cp = CooPlan(A)
        if cp.info()["sorted"]:
            return cls(sparse_CSR_t(cp.row_ptr(), A.col, A.val, A.N, A.nnz), tile_rows, panel_cols)
        from .convert import coo_csr
        return cls(coo_csr(A), tile_rows, panel_cols)

    def __del__(self):
        try:
            if self.handle:
                _lib.load().sable_csr_tiles_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class CooPlan:
    """Row analysis of a COO matrix (sable_coo_plan_*): rows checked once;
    row-sorted input keeps its row heads and row blocks, so a product is
    launch-only."""

    def __init__(self, A):
        self.ctx = A.row.ctx
        self.A = A
        h = c_vp()
        call("sable_coo_plan_create", self.ctx.handle, A.N, A.nnz, A.row.ptr, ctypes.byref(h))
        self.handle = h

    def info(self):
        srt, nb = ctypes.c_int(), c_i64()
        call("sable_coo_plan_info", self.handle, ctypes.byref(srt), ctypes.byref(nb))
        return {"sorted": bool(srt.value), "nblocks": nb.value}

    def row_ptr(self):
        """A copy of the row heads of a row-sorted COO (raises otherwise)."""
        from .core import DeviceArray
        p = c_vp()
        call("sable_coo_plan_row_ptr", self.handle, ctypes.byref(p))
        d = DeviceArray(self.A.N + 1, np.int32, self.ctx)
        call("sable_memcpy_d2d", self.ctx.handle, d.ptr, p.value, d.nbytes)
        return d

    def run(self, x, y, beta=1.0, inner_max=1):
        A = self.A
        call("sable_spmv_coo_plan", self.ctx.handle, self.handle, dtype_code(A.dtype), A.row.ptr, A.col.ptr,
             A.val.ptr, x.ptr, y.ptr, float(beta), int(inner_max))

    def __del__(self):
        try:
            if self.handle:
                _lib.load().sable_coo_plan_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class CsrAuto:
    """Autotuned CSR schedule of a resident matrix (sable_csr_auto_*): the
    candidates are timed on the device and the fastest kept; the choice is
    cached per matrix class.  deterministic=True keeps to the bit-
    reproducible schedules."""

    NAMES = {0: "blocks", 1: "vector", 2: "tiles"}

    def __init__(self, A, deterministic=False, timing=True, ncols=0):
        self.ctx = A.row.ctx
        self.A = A
        h = c_vp()
        flags = (1 if deterministic else 0) | (0 if timing else 2)
        call("sable_csr_auto_create", self.ctx.handle, dtype_code(A.dtype), A.N, int(ncols), A.row.ptr, A.col.ptr,
             A.val.ptr, flags, ctypes.byref(h))
        self.handle = h

    def info(self):
        i = _lib.CsrAutoInfo()
        call("sable_csr_auto_info", self.handle, ctypes.byref(i))
        return {"schedule": self.NAMES.get(i.schedule, i.schedule), "param0": i.param0, "param1": i.param1,
                "ncols": i.ncols, "nnz": i.nnz, "extra_bytes": i.extra_bytes, "ms": i.ms,
                "from_cache": bool(i.from_cache)}

    def run(self, x, y, beta=1.0, inner_max=1):
        call("sable_spmv_csr_auto", self.ctx.handle, self.handle, x.ptr, y.ptr, float(beta), int(inner_max))

    @staticmethod
    def clear_cache():
        _lib.load().sable_csr_auto_cache_clear()

    def __del__(self):
        try:
            if self.handle:
                _lib.load().sable_csr_auto_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class CsrPanels:
    """Column-panel blocked CSR plan (sable_csr_panels_*): a panel-major copy
    of the matrix, one launch per panel of panel_cols columns."""

    def __init__(self, A, panel_cols=0, tune=None):
        self.ctx = A.row.ctx
        self.A = A
        h = c_vp()
        t = None
        if tune is not None:
            t = ctypes.byref(_lib.CsrTune(int(tune.get("block_weight", 0)), int(tune.get("row_weight", 0))))
        call("sable_csr_panels_create", self.ctx.handle, dtype_code(A.dtype), A.N, A.row.ptr, A.col.ptr, A.val.ptr,
             int(panel_cols), t, ctypes.byref(h))
        self.handle = h

    def info(self):
        n, w, b = c_i32(), c_i32(), c_i64()
        call("sable_csr_panels_info", self.handle, ctypes.byref(n), ctypes.byref(w), ctypes.byref(b))
        return {"npanels": n.value, "panel_cols": w.value, "bytes": b.value}

    def run(self, x, y, beta=1.0, inner_max=1):
        call("sable_spmv_csr_panels", self.ctx.handle, self.handle, x.ptr, y.ptr, float(beta), int(inner_max))

    def __del__(self):
        try:
            if self.handle:
                _lib.load().sable_csr_panels_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def _check_vec(A, x, y):
    if x.vec.dtype != A.dtype or y.vec.dtype != A.dtype:
        raise TypeError("x, y must have the matrix dtype %s" % A.dtype)
    if x.vec_nelem < A.N or y.vec_nelem < A.N:
        raise ValueError("x and y need at least N=%d elements" % A.N)


def spmv(A, x_view, y_view, inner_max=1, beta=1.0, use_plan=True):
    """y = A.x + beta*y, inner_max times (beta=1: the reference's y += A.x)."""
    if A is None:
        print("spmv input : undefined", file=sys.stderr)
        return
    _check_vec(A, x_view, y_view)
    ctx = x_view.vec.ctx.handle
    dt = dtype_code(A.dtype)
    x, y = x_view.vec.ptr, y_view.vec.ptr
    if isinstance(A, sparse_COO_t):
        if use_plan:
            A.plan().run(x_view.vec, y_view.vec, beta, inner_max)
        else:
            call("sable_spmv_coo", ctx, dt, A.N, A.nnz, A.row.ptr, A.col.ptr, A.val.ptr, x, y, float(beta),
                 int(inner_max))
    elif isinstance(A, sparse_CSR_t):
        if use_plan:
            A.plan().run(x_view.vec, y_view.vec, beta, inner_max)
        else:
            call("sable_spmv_csr", ctx, dt, A.N, A.row.ptr, A.col.ptr, A.val.ptr, x, y, float(beta),
                 int(inner_max))
    elif isinstance(A, sparse_DIA_t):
        call("sable_spmv_dia", ctx, dt, A.N, A.ndiags, A.stride, A.offset.ptr, A.data.ptr, x, y, float(beta),
             int(inner_max))
    elif isinstance(A, sparse_ELL_t):
        call("sable_spmv_ell", ctx, dt, A.N, A.ncols, A.indices.ptr, A.data.ptr, x, y, float(beta),
             int(inner_max))
    else:
        raise TypeError("unsupported matrix type %r" % type(A))
