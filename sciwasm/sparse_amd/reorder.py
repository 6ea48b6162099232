"""Row reordering of resident CSR matrices (par/src/sparse32.js:2148-2235).

``sort_rows_by_nnz(A)`` returns the CSR with rows in ascending nnz order
(stable) and the fields the reference attaches to it: ``permutation``
(new row i = old row permutation[i]), ``nnz_row`` (nnz of each old row) and
the short-row counts ``num_zero_rows`` ... ``num_three_rows`` (cumulative).
``sort_y_rows_by_nnz(y, A)`` puts a result of the reordered matrix back in
the original row order.
"""
import ctypes

import numpy as np

from ._lib import call, c_i32
from .core import DeviceArray, sparse_CSR_t, dtype_code


def sort_rows_by_nnz(A):
    if not isinstance(A, sparse_CSR_t):
        raise TypeError("sort_rows_by_nnz takes a CSR matrix")
    ctx = A.row.ctx
    N = A.N
    rp = DeviceArray(N + 1, np.int32, ctx)
    col = DeviceArray(A.nnz, np.int32, ctx)
    val = DeviceArray(A.nnz, A.dtype, ctx)
    perm = DeviceArray(N, np.int32, ctx)
    nnz_row = DeviceArray(N, np.int32, ctx)
    bins = (c_i32 * 4)()
    call("sable_csr_sort_rows_by_nnz", ctx.handle, dtype_code(A.dtype), N, A.row.ptr, A.col.ptr, A.val.ptr, rp.ptr,
         col.ptr, val.ptr, perm.ptr, nnz_row.ptr, bins)
    B = sparse_CSR_t(rp, col, val, N, A.nnz)
    B.permutation = perm
    B.nnz_row = nnz_row
    B.num_zero_rows, B.num_one_rows, B.num_two_rows, B.num_three_rows = (int(b) for b in bins)
    return B


def reorder_NN(A, w=16):
    """par/src/sparse32.js:2501-2622: locality ordering of a matrix returned
    by sort_rows_by_nnz (rows of equal nnz regrouped so that consecutive rows
    share cache lines of x).  The result carries the composed permutation and
    the short-row counts of A."""
    if not isinstance(A, sparse_CSR_t):
        raise TypeError("reorder_NN takes a CSR matrix")
    ctx = A.row.ctx
    N = A.N
    rp = DeviceArray(N + 1, np.int32, ctx)
    col = DeviceArray(A.nnz, np.int32, ctx)
    val = DeviceArray(A.nnz, A.dtype, ctx)
    perm = DeviceArray(N, np.int32, ctx)
    pin = getattr(A, "permutation", None)
    call("sable_csr_reorder_nn", ctx.handle, dtype_code(A.dtype), N, A.row.ptr, A.col.ptr, A.val.ptr,
         pin.ptr if pin is not None else None, int(w), rp.ptr, col.ptr, val.ptr, perm.ptr)
    B = sparse_CSR_t(rp, col, val, N, A.nnz)
    B.permutation = perm
    for k in ("nnz_row", "num_zero_rows", "num_one_rows", "num_two_rows", "num_three_rows"):
        if hasattr(A, k):
            setattr(B, k, getattr(A, k))
    return B


def sort_y_rows_by_nnz(y_view, A):
    """y_new[A.permutation[i]] = y[i]; y_view then holds y_new (the
    reference swaps y_view's storage the same way)."""
    d = y_view.vec
    out = DeviceArray(d.n, d.dtype, d.ctx)
    call("sable_permute_rows", d.ctx.handle, dtype_code(d.dtype), A.N, A.permutation.ptr, 1, d.ptr, out.ptr)
    y_view.vec = out
    return y_view


def permute_rows(x, perm, scatter=False):
    """out[i] = x[perm[i]] (or out[perm[i]] = x[i] with scatter=True) for a
    DeviceArray x."""
    out = DeviceArray(x.n, x.dtype, x.ctx)
    call("sable_permute_rows", x.ctx.handle, dtype_code(x.dtype), perm.n, perm.ptr, int(scatter), x.ptr, out.ptr)
    return out
