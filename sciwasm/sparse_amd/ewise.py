"""Element-wise operations and eliminate_zeros on resident matrices.

Out-of-place functions mirror seq/src/sparse32.js:671-1207 (``expm1(A)``
... ``nearest(A)``, ``pow(A, p)``, ``eliminate_zeros(A)``): each returns a
new matrix of the same format.  The in-place methods of the reference's
objects (seq/src/sparse32.js:200-660, ``A.expm1()``, ``A.pow(p)``,
``A.multiply(s)``, ``A.eliminate_zeros()``, ``A.transpose()``) are attached
to the matrix classes here.

Per-format rules, as in the reference's WebAssembly exports:
* COO and CSR: every stored value; ``sign`` is ``v > 0 ? 1 : -1``.
* DIA: only the in-range slots of each diagonal; ``sign`` keeps 0.
* ELL: every slot, padding included; ``sign`` keeps 0.
"""
import ctypes
import math

import numpy as np

from ._lib import call, c_i32, c_i64
from .core import (DeviceArray, sparse_COO_t, sparse_CSR_t, sparse_DIA_t, sparse_ELL_t, dtype_code)

OPS = ["expm1", "log1p", "sin", "tan", "pow", "deg2rad", "rad2deg", "sign", "sign0", "abs", "neg", "sqrt", "ceil",
       "floor", "trunc", "nearest", "multiply"]
UNARY = ["expm1", "log1p", "sin", "tan", "deg2rad", "rad2deg", "sign", "abs", "neg", "sqrt", "ceil", "floor",
         "trunc", "nearest"]


def _code(op, fmt):
    if op == "sign" and fmt in ("dia", "ell"):
        op = "sign0"
    return OPS.index(op)


def _param(op, p):
    if op in ("deg2rad", "rad2deg"):
        return math.pi  # the reference passes Math.PI
    return float(p)


def _values(ctx, dtype, code, param, n, src, dst):
    call("sable_ew_values", ctx.handle, dtype_code(dtype), code, ctypes.c_double(param), n, src.ptr, dst.ptr)


def _apply(A, op, p=0.0, inplace=False):
    param = _param(op, p)
    if isinstance(A, (sparse_COO_t, sparse_CSR_t)):
        ctx = A.val.ctx
        out = A.val if inplace else DeviceArray(A.nnz, A.dtype, ctx)
        _values(ctx, A.dtype, _code(op, "coo"), param, A.nnz, A.val, out)
        if inplace:
            return A
        if isinstance(A, sparse_COO_t):
            return sparse_COO_t(A.row.copy(), A.col.copy(), out, A.N, A.nnz)
        return sparse_CSR_t(A.row.copy(), A.col.copy(), out, A.N, A.nnz)
    if isinstance(A, sparse_DIA_t):
        ctx = A.data.ctx
        out = A.data if inplace else DeviceArray.zeros(A.data.n, A.dtype, ctx)
        call("sable_ew_dia", ctx.handle, dtype_code(A.dtype), _code(op, "dia"), ctypes.c_double(param), A.N,
             A.ndiags, A.stride, A.offset.ptr, 0, A.N, A.data.ptr, out.ptr)
        if inplace:
            return A
        return sparse_DIA_t(A.offset.copy(), out, A.ndiags, A.stride, A.N, A.nnz)
    if isinstance(A, sparse_ELL_t):
        ctx = A.data.ctx
        out = A.data if inplace else DeviceArray(A.data.n, A.dtype, ctx)
        _values(ctx, A.dtype, _code(op, "ell"), param, A.ncols * A.N, A.data, out)
        if inplace:
            return A
        return sparse_ELL_t(A.indices.copy(), out, A.ncols, A.N, A.nnz)
    raise TypeError("element-wise ops take COO, CSR, DIA or ELL matrices")


def _make(op):
    def f(A):
        return _apply(A, op)
    f.__name__ = op
    f.__doc__ = "seq/src/sparse32.js %s(A_in): a new matrix with %s applied to the stored values." % (op, op)
    return f


expm1, log1p, sin, tan, deg2rad, rad2deg, sign, abs, neg, sqrt, ceil, floor, trunc, nearest = \
    [_make(op) for op in UNARY]


def pow(A, p):
    """seq/src/sparse32.js:787-814: Math.pow(v, p) on every stored value."""
    return _apply(A, "pow", p)


# ------------------------------------------------------------------ eliminate_zeros
def count_nonzeros(A):
    """Stored values with v > 0 || v < 0 (count_nonzeros_coo / _dia / _ell).
    DIA returns (count_diags, count); ELL returns (count_cols, count)."""
    if isinstance(A, (sparse_COO_t, sparse_CSR_t)):
        c = c_i64()
        call("sable_count_nonzeros", A.val.ctx.handle, dtype_code(A.dtype), A.nnz, A.val.ptr, ctypes.byref(c))
        return c.value
    if isinstance(A, sparse_DIA_t):
        cd, c = c_i32(), c_i64()
        call("sable_count_nonzeros_dia", A.data.ctx.handle, dtype_code(A.dtype), A.N, A.ndiags, A.stride,
             A.offset.ptr, A.data.ptr, ctypes.byref(cd), ctypes.byref(c))
        return cd.value, c.value
    if isinstance(A, sparse_ELL_t):
        cc, c = c_i32(), c_i64()
        call("sable_count_nonzeros_ell", A.data.ctx.handle, dtype_code(A.dtype), A.N, A.ncols, A.data.ptr,
             ctypes.byref(cc), ctypes.byref(c))
        return cc.value, c.value
    raise TypeError("count_nonzeros takes COO, CSR, DIA or ELL matrices")


def _copy(A):
    if isinstance(A, sparse_COO_t):
        return sparse_COO_t(A.row.copy(), A.col.copy(), A.val.copy(), A.N, A.nnz)
    if isinstance(A, sparse_CSR_t):
        return sparse_CSR_t(A.row.copy(), A.col.copy(), A.val.copy(), A.N, A.nnz)
    if isinstance(A, sparse_DIA_t):
        return sparse_DIA_t(A.offset.copy(), A.data.copy(), A.ndiags, A.stride, A.N, A.nnz)
    return sparse_ELL_t(A.indices.copy(), A.data.copy(), A.ncols, A.N, A.nnz)


def _compacted(A):
    """The matrix without zeros, always computed (copy_nonzeros_*)."""
    if isinstance(A, sparse_COO_t):
        ctx = A.val.ctx
        c = count_nonzeros(A)
        r, co, v = DeviceArray(c, np.int32, ctx), DeviceArray(c, np.int32, ctx), DeviceArray(c, A.dtype, ctx)
        got = c_i64()
        call("sable_eliminate_zeros_coo", ctx.handle, dtype_code(A.dtype), A.nnz, A.row.ptr, A.col.ptr, A.val.ptr,
             r.ptr, co.ptr, v.ptr, ctypes.byref(got))
        return sparse_COO_t(r, co, v, A.N, c)
    if isinstance(A, sparse_CSR_t):
        ctx = A.val.ctx
        c = count_nonzeros(A)
        rp, co, v = DeviceArray(A.N + 1, np.int32, ctx), DeviceArray(c, np.int32, ctx), DeviceArray(c, A.dtype, ctx)
        got = c_i64()
        call("sable_eliminate_zeros_csr", ctx.handle, dtype_code(A.dtype), A.N, A.row.ptr, A.col.ptr, A.val.ptr,
             rp.ptr, co.ptr, v.ptr, ctypes.byref(got))
        return sparse_CSR_t(rp, co, v, A.N, c)
    if isinstance(A, sparse_DIA_t):
        ctx = A.data.ctx
        cd, c = count_nonzeros(A)
        off, data = DeviceArray(cd, np.int32, ctx), DeviceArray(cd * A.stride, A.dtype, ctx)
        got = c_i32()
        call("sable_eliminate_zeros_dia", ctx.handle, dtype_code(A.dtype), A.N, A.ndiags, A.stride, A.offset.ptr,
             A.data.ptr, off.ptr, data.ptr, ctypes.byref(got))
        return sparse_DIA_t(off, data, cd, A.stride, A.N, c)
    if isinstance(A, sparse_ELL_t):
        ctx = A.data.ctx
        cc, c = count_nonzeros(A)
        data, idx = DeviceArray(cc * A.N, A.dtype, ctx), DeviceArray(cc * A.N, np.int32, ctx)
        call("sable_eliminate_zeros_ell", ctx.handle, dtype_code(A.dtype), A.N, A.ncols, cc, A.data.ptr,
             A.indices.ptr, data.ptr, idx.ptr)
        return sparse_ELL_t(idx, data, cc, A.N, c)
    raise TypeError("eliminate_zeros takes COO, CSR, DIA or ELL matrices")


def _same_count(A):
    c = count_nonzeros(A)
    return (c if isinstance(A, (sparse_COO_t, sparse_CSR_t)) else c[1]) == A.nnz


def eliminate_zeros(A):
    """seq/src/sparse32.js:1140-1207: a new matrix without explicit zeros.
    When the nonzero count equals A.nnz the reference copies A unchanged."""
    return _copy(A) if _same_count(A) else _compacted(A)


def _self_eliminate_zeros(self):
    """seq/src/sparse32.js:298-314 and siblings: in place; nothing happens
    when the nonzero count equals nnz."""
    if _same_count(self):
        return self
    B = _compacted(self)
    self.__dict__.update({k: v for k, v in B.__dict__.items()})
    return self


def _self_transpose(self):
    """seq/src/sparse32.js:285-295 (COO: rows and columns swapped, then
    sorted by (row, col)), :405-420 (CSR), :510-527 (DIA), :626-641 (ELL)."""
    from .convert import transpose
    if isinstance(self, sparse_COO_t):
        # rows and columns swapped, then sorted by (row, col) on the device
        self.row, self.col = self.col, self.row
        call("sable_coo_sort", self.val.ctx.handle, dtype_code(self.dtype), self.N, self.nnz, self.row.ptr,
             self.col.ptr, self.val.ptr)
        self.invalidate()
        return self
    B = transpose(self)
    self.__dict__.update({k: v for k, v in B.__dict__.items()})
    return self


def _attach():
    def method(op):
        def m(self):
            return _apply(self, op, inplace=True)
        m.__name__ = op
        return m

    for cls in (sparse_COO_t, sparse_CSR_t, sparse_DIA_t, sparse_ELL_t):
        for op in UNARY:
            setattr(cls, op, method(op))
        cls.pow = lambda self, p: _apply(self, "pow", p, inplace=True)
        cls.eliminate_zeros = _self_eliminate_zeros
        cls.transpose = _self_transpose
    sparse_COO_t.multiply = lambda self, s: _apply(self, "multiply", s, inplace=True)


_attach()
