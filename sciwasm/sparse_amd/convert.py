"""Format conversions and transposes, mirroring seq/src/sparse32.js.

The storage guards are applied here, where the reference applies them (in
its JS): csr_ell and csr_dia print the reference's message and return None
when the padded storage would reach 2^27 slots or exceed 5x nnz
(seq/src/sparse32.js:1457-1463, 1525-1531); the message goes to stderr so
that stdout stays free for the caller's data.  ``guard=False`` lifts the
2^27 limit, which only exists because of the reference's 2 GiB wasm32
memory; the 5x ratio check stays.
"""
import ctypes
import sys

import numpy as np

from ._lib import call, c_i32
from .core import (DeviceArray, sparse_COO_t, sparse_CSC_t, sparse_CSR_t, sparse_DIA_t, sparse_ELL_t,
                   dtype_code)

GUARD_SLOTS = 2 ** 27
GUARD_RATIO = 5


def coo_csr(A):
    """seq/src/sparse32.js:1674-1720 on the GPU (bit-identical)."""
    ctx = A.row.ctx
    rp = DeviceArray(A.N + 1, np.int32, ctx)
    col = DeviceArray(A.nnz, np.int32, ctx)
    val = DeviceArray(A.nnz, A.dtype, ctx)
    call("sable_coo_csr", ctx.handle, dtype_code(A.dtype), A.N, A.nnz, A.row.ptr, A.col.ptr, A.val.ptr,
         rp.ptr, col.ptr, val.ptr)
    return sparse_CSR_t(rp, col, val, A.N, A.nnz)


def coo_csc(A):
    """seq/src/sparse32.js:1629-1671 on the GPU (bit-identical): CSC with
    rows ascending in each column, duplicates in input order."""
    ctx = A.row.ctx
    cp = DeviceArray(A.N + 1, np.int32, ctx)
    row = DeviceArray(A.nnz, np.int32, ctx)
    val = DeviceArray(A.nnz, A.dtype, ctx)
    call("sable_coo_csc", ctx.handle, dtype_code(A.dtype), A.N, A.nnz, A.row.ptr, A.col.ptr, A.val.ptr,
         cp.ptr, row.ptr, val.ptr)
    return sparse_CSC_t(cp, row, val, A.N, A.nnz)


def num_cols(A):
    """seq/src/sparse32.js:1436-1447."""
    nc = c_i32()
    call("sable_csr_num_cols", A.row.ctx.handle, A.N, A.row.ptr, ctypes.byref(nc))
    return nc.value


def _allowed(slots, nnz, guard):
    if guard and not slots < GUARD_SLOTS:
        return False
    return nnz > 0 and slots / nnz <= GUARD_RATIO


def csr_ell(A, guard=True):
    """seq/src/sparse32.js:1449-1484; None when the guard refuses."""
    nc = num_cols(A)
    if not _allowed(nc * A.N, A.nnz, guard):
        print("not efficient to store this matrix in ELL", file=sys.stderr)
        return None
    ctx = A.row.ctx
    idx = DeviceArray(nc * A.N, np.int32, ctx)
    data = DeviceArray(nc * A.N, A.dtype, ctx)
    call("sable_csr_ell", ctx.handle, dtype_code(A.dtype), A.N, nc, A.row.ptr, A.col.ptr, A.val.ptr, idx.ptr,
         data.ptr)
    E = sparse_ELL_t(idx, data, nc, A.N, A.nnz)
    E.major = "col"
    return E


def num_diags(A, stride_mode="N"):
    """seq/src/sparse32.js:1486-1512 -> [nd, stride].  stride_mode "N-min"
    follows seq/src/sparse64.js:1066-1092 (stride = N - min|offset|)."""
    nd, mn = c_i32(), c_i32()
    call("sable_csr_num_diags", A.row.ctx.handle, A.N, A.row.ptr, A.col.ptr, ctypes.byref(nd), ctypes.byref(mn))
    stride = A.N if stride_mode == "N" else A.N - mn.value
    return [nd.value, stride]


def csr_dia(A, guard=True, stride_mode="N"):
    """seq/src/sparse32.js:1515-1574; None when the guard refuses."""
    nd, stride = num_diags(A, stride_mode)
    if not _allowed(nd * stride, A.nnz, guard):
        print("not efficient to store this matrix in DIA", file=sys.stderr)
        return None
    ctx = A.row.ctx
    off = DeviceArray(max(nd, 1), np.int32, ctx)
    data = DeviceArray(max(nd * stride, 1), A.dtype, ctx)
    call("sable_csr_dia", ctx.handle, dtype_code(A.dtype), A.N, nd, stride, A.row.ptr, A.col.ptr, A.val.ptr,
         off.ptr, data.ptr)
    D = sparse_DIA_t(off, data, nd, stride, A.N, A.nnz)
    D.major = "col"
    return D


def transpose(A):
    """seq/src/sparse32.js:1106-1138 for every format."""
    ctx = (A.row if isinstance(A, (sparse_COO_t, sparse_CSR_t)) else A.data).ctx
    dt = dtype_code(A.dtype)
    if isinstance(A, sparse_COO_t):
        # a copy with row and column swapped, no sort (seq/src/sparse32.js:1109-1114)
        return sparse_COO_t(A.col.copy(), A.row.copy(), A.val.copy(), A.N, A.nnz)
    if isinstance(A, sparse_CSR_t):
        rp = DeviceArray(A.N + 1, np.int32, ctx)
        col = DeviceArray(A.nnz, np.int32, ctx)
        val = DeviceArray(A.nnz, A.dtype, ctx)
        call("sable_transpose_csr", ctx.handle, dt, A.N, A.row.ptr, A.col.ptr, A.val.ptr, rp.ptr, col.ptr, val.ptr)
        return sparse_CSR_t(rp, col, val, A.N, A.nnz)
    if isinstance(A, sparse_DIA_t):
        off = DeviceArray(max(A.ndiags, 1), np.int32, ctx)
        data = DeviceArray(max(A.ndiags * A.stride, 1), A.dtype, ctx)
        call("sable_transpose_dia", ctx.handle, dt, A.N, A.ndiags, A.stride, A.offset.ptr, A.data.ptr, off.ptr,
             data.ptr)
        return sparse_DIA_t(off, data, A.ndiags, A.stride, A.N, A.nnz)
    if isinstance(A, sparse_ELL_t):
        nco = c_i32()
        call("sable_count_transpose_cols_ell", ctx.handle, dt, A.N, A.ncols, A.data.ptr, A.indices.ptr,
             ctypes.byref(nco))
        idx = DeviceArray(max(nco.value * A.N, 1), np.int32, ctx)
        data = DeviceArray(max(nco.value * A.N, 1), A.dtype, ctx)
        call("sable_transpose_ell", ctx.handle, dt, A.N, A.ncols, nco.value, A.data.ptr, A.indices.ptr, data.ptr,
             idx.ptr)
        return sparse_ELL_t(idx, data, nco.value, A.N, A.nnz)
    raise TypeError("unsupported matrix type %r" % type(A))
