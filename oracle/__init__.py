"""CPU parity oracle for the SciWasm.Sparse Level-II hot path.

TEST INFRASTRUCTURE ONLY: this package may be imported by ``tests/``, by
``__graft_entry__.smoke()`` and by ``bench.py``'s ``cpu_baseline`` leg, and
there only as the checker (or as the timed CPU baseline).  The product
package ``sciwasm.sparse_amd`` never imports it.

It wraps ``liboracle.so`` (built from ``sable_oracle.c`` by ``make -C
oracle``), a C restatement of the reference routines in their own operation
order; each C function cites the reference file:line it follows.  The
wrappers below take and return numpy arrays.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_F32P = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_F64P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_VP = ctypes.c_void_p
_I = ctypes.c_int


def build():
    """Compile liboracle.so in place (gcc, a few seconds)."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        for suf, FP in (("f32", _F32P), ("f64", _F64P)):
            getattr(L, "oracle_coo_csr_" + suf).argtypes = [_I, _I, _I32P, _I32P, FP, _I32P, _I32P, FP]
            getattr(L, "oracle_csr_ell_" + suf).argtypes = [_I, _I, _I32P, _I32P, FP, _I32P, FP]
            getattr(L, "oracle_csr_dia_" + suf).argtypes = [_I, _I, _I, _I32P, _I32P, FP, _I32P, FP]
            getattr(L, "oracle_count_transpose_cols_ell_" + suf).argtypes = [_I, _I, FP, _I32P]
            getattr(L, "oracle_count_transpose_cols_ell_" + suf).restype = _I
            getattr(L, "oracle_transpose_ell_" + suf).argtypes = [_I, _I, _I, FP, _I32P, FP, _I32P]
            getattr(L, "oracle_level_reorder_" + suf).argtypes = [_I, _I32P, _I32P, FP, _I32P, _I32P, _I32P, _I32P, FP]
            getattr(L, "oracle_spmv_coo_" + suf).argtypes = [_I, _I32P, _I32P, FP, FP, FP, _I]
            getattr(L, "oracle_spmv_csr_" + suf).argtypes = [_I, _I32P, _I32P, FP, FP, FP, _I]
            getattr(L, "oracle_spmv_dia_" + suf).argtypes = [_I, _I, _I, _I32P, FP, FP, FP, _I]
            getattr(L, "oracle_spmv_ell_" + suf).argtypes = [_I, _I, _I32P, FP, FP, FP, _I]
            getattr(L, "oracle_csr_ell_row_" + suf).argtypes = [_I, _I, _I32P, _I32P, FP, _I32P, FP]
            getattr(L, "oracle_csr_dia_row_" + suf).argtypes = [_I, _I, _I32P, _I32P, FP, _I32P, FP]
            getattr(L, "oracle_spmv_dia_row_" + suf).argtypes = [_I, _I, _I32P, FP, _I, _I, FP, FP, _I]
            getattr(L, "oracle_spmv_ell_row_" + suf).argtypes = [_I, _I, _I32P, FP, _I, _I, FP, FP, _I, _I]
            getattr(L, "oracle_spts_level_" + suf).argtypes = [_I, _I32P, _I32P, FP, FP, FP]
            getattr(L, "oracle_spts_syncfree_" + suf).argtypes = [_I, _I32P, _I32P, FP, FP, FP]
            getattr(L, "oracle_spts_csc_" + suf).argtypes = [_I, _I32P, _I32P, FP, FP, FP]
            getattr(L, "oracle_fletcher_" + suf).argtypes = [_I, FP]
            getattr(L, "oracle_fletcher_" + suf).restype = ctypes.c_double
        L.oracle_num_cols.argtypes = [_I, _I32P]
        L.oracle_num_cols.restype = _I
        L.oracle_num_diags.argtypes = [_I, _I32P, _I32P, ctypes.POINTER(_I)]
        L.oracle_num_diags.restype = _I
        L.oracle_transpose_csr.argtypes = [_I, _I, _I, _I32P, _I32P, _VP, _I32P, _I32P, _VP]
        L.oracle_transpose_dia.argtypes = [_I, _I, _I, _I, _I32P, _VP, _I32P, _VP]
        L.oracle_level_sets.argtypes = [_I, _I32P, _I32P, _I32P, _I32P, _I32P, _I32P]
        L.oracle_level_sets.restype = _I
        L.oracle_spmv_csr_ld.argtypes = [_I, _I32P, _I32P, _F64P, _F64P, _VP, _F64P, _VP]
        L.oracle_spts_ld.argtypes = [_I, _I32P, _I32P, _F64P, _F64P, _F64P]
        L.oracle_rel_residual.argtypes = [_I, _I32P, _I32P, _F64P, _F64P, _F64P]
        L.oracle_rel_residual.restype = ctypes.c_double
        L.oracle_time_spmv_csr_f64.argtypes = [_I, _I32P, _I32P, _F64P, _F64P, _F64P, _I, _I]
        L.oracle_time_spmv_csr_f64.restype = ctypes.c_double
        L.oracle_time_spmv_csr_f64_mt.argtypes = [_I, _I32P, _I32P, _F64P, _F64P, _F64P, _I, _I, _I]
        L.oracle_time_spmv_csr_f64_mt.restype = ctypes.c_double
        L.oracle_static_nnz.argtypes = [_I, _I32P, _I, _I32P, _I32P]
        for suf in ("f32", "f64"):
            fn = getattr(L, "oracle_mm_coo_" + suf)
            fn.argtypes = [ctypes.c_char_p, ctypes.c_size_t, _I, ctypes.POINTER(ctypes.c_int64), _VP, _VP, _VP]
            fn.restype = _I
        _LIB = L
    return _LIB


def _suf(a):
    if a.dtype == np.float32:
        return "f32"
    if a.dtype == np.float64:
        return "f64"
    raise TypeError("values must be float32 or float64, got %s" % a.dtype)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------- conversions
def coo_csr(N, row, col, val):
    """seq/src/sparse32.js:1674-1720 -> (rowptr, col, val)."""
    row, col, val = _i32(row), _i32(col), np.ascontiguousarray(val)
    nnz = len(row)
    rp = np.zeros(N + 1, np.int32)
    oc = np.zeros(nnz, np.int32)
    ov = np.zeros(nnz, val.dtype)
    getattr(lib(), "oracle_coo_csr_" + _suf(val))(N, nnz, row, col, val, rp, oc, ov)
    return rp, oc, ov


def num_cols(N, rowptr):
    return lib().oracle_num_cols(N, _i32(rowptr))


def num_diags(N, rowptr, col):
    """Returns (nd, min_abs_offset)."""
    mn = _I(0)
    nd = lib().oracle_num_diags(N, _i32(rowptr), _i32(col), ctypes.byref(mn))
    return nd, mn.value


def csr_ell(N, rowptr, col, val, nc=None):
    rowptr, col, val = _i32(rowptr), _i32(col), np.ascontiguousarray(val)
    if nc is None:
        nc = num_cols(N, rowptr)
    idx = np.zeros(nc * N, np.int32)
    data = np.zeros(nc * N, val.dtype)
    getattr(lib(), "oracle_csr_ell_" + _suf(val))(N, nc, rowptr, col, val, idx, data)
    return nc, idx, data


def csr_dia(N, rowptr, col, val, stride=None):
    """Returns (nd, stride, offset, data); stride defaults to N."""
    rowptr, col, val = _i32(rowptr), _i32(col), np.ascontiguousarray(val)
    nd, _ = num_diags(N, rowptr, col)
    if stride is None:
        stride = N
    off = np.zeros(max(nd, 1), np.int32)
    data = np.zeros(max(nd * stride, 1), val.dtype)
    getattr(lib(), "oracle_csr_dia_" + _suf(val))(N, nd, stride, rowptr, col, val, off, data)
    return nd, stride, off[:nd], data[:nd * stride]


def transpose_csr(N, rowptr, col, val):
    rowptr, col, val = _i32(rowptr), _i32(col), np.ascontiguousarray(val)
    nnz = len(col)
    orp = np.zeros(N + 1, np.int32)
    oc = np.zeros(max(nnz, 1), np.int32)
    ov = np.zeros(max(nnz, 1), val.dtype)
    lib().oracle_transpose_csr(N, nnz, val.itemsize, rowptr, col, val.ctypes.data,
                               orp, oc, ov.ctypes.data)
    return orp, oc[:nnz], ov[:nnz]


def transpose_dia(N, stride, offset, data):
    offset, data = _i32(offset), np.ascontiguousarray(data)
    nd = len(offset)
    oo = np.zeros(max(nd, 1), np.int32)
    od = np.zeros(max(nd * stride, 1), data.dtype)
    lib().oracle_transpose_dia(N, nd, stride, data.itemsize, offset, data.ctypes.data,
                               oo, od.ctypes.data)
    return oo[:nd], od[:nd * stride]


def transpose_ell(N, nc, indices, data):
    """Returns (nc_out, indices_out, data_out)."""
    indices, data = _i32(indices), np.ascontiguousarray(data)
    s = _suf(data)
    nco = getattr(lib(), "oracle_count_transpose_cols_ell_" + s)(N, nc, data, indices)
    oi = np.zeros(max(nco * N, 1), np.int32)
    od = np.zeros(max(nco * N, 1), data.dtype)
    getattr(lib(), "oracle_transpose_ell_" + s)(N, nc, nco, data, indices, od, oi)
    return nco, oi[:nco * N], od[:nco * N]


# ------------------------------------------------------------------ level sets
def level_sets(N, rowptr, col):
    """par/src/sparse32.js:2942-3181 -> (nlevels, level_ptr, permutation, order, level_of_row)."""
    rowptr, col = _i32(rowptr), _i32(col)
    lvl = np.zeros(max(N, 1), np.int32)
    lp = np.zeros(N + 2, np.int32)
    perm = np.zeros(max(N, 1), np.int32)
    order = np.zeros(max(N, 1), np.int32)
    nl = lib().oracle_level_sets(N, rowptr, col, lvl, lp, perm, order)
    return nl, lp[:nl + 1].copy(), perm[:N], order[:N], lvl[:N]


def level_reorder(N, rowptr, col, val, permutation, order):
    rowptr, col, val = _i32(rowptr), _i32(col), np.ascontiguousarray(val)
    nrp = np.zeros(N + 1, np.int32)
    nc = np.zeros(max(len(col), 1), np.int32)
    nv = np.zeros(max(len(col), 1), val.dtype)
    getattr(lib(), "oracle_level_reorder_" + _suf(val))(N, rowptr, col, val, _i32(permutation),
                                                        _i32(order), nrp, nc, nv)
    return nrp, nc[:len(col)], nv[:len(col)]


# ------------------------------------------------------------------------ spmv
def spmv_coo(row, col, val, x, y, inner_max=1):
    y = np.array(y, dtype=val.dtype, copy=True)
    getattr(lib(), "oracle_spmv_coo_" + _suf(val))(len(row), _i32(row), _i32(col),
                                                  np.ascontiguousarray(val),
                                                  np.ascontiguousarray(x, val.dtype), y, inner_max)
    return y


def spmv_csr(N, rowptr, col, val, x, y, inner_max=1):
    y = np.array(y, dtype=val.dtype, copy=True)
    getattr(lib(), "oracle_spmv_csr_" + _suf(val))(N, _i32(rowptr), _i32(col),
                                                  np.ascontiguousarray(val),
                                                  np.ascontiguousarray(x, val.dtype), y, inner_max)
    return y


def spmv_dia(N, stride, offset, data, x, y, inner_max=1):
    y = np.array(y, dtype=data.dtype, copy=True)
    getattr(lib(), "oracle_spmv_dia_" + _suf(data))(N, len(offset), stride, _i32(offset),
                                                   np.ascontiguousarray(data),
                                                   np.ascontiguousarray(x, data.dtype), y, inner_max)
    return y


def spmv_ell(N, nc, indices, data, x, y, inner_max=1):
    y = np.array(y, dtype=data.dtype, copy=True)
    getattr(lib(), "oracle_spmv_ell_" + _suf(data))(N, nc, _i32(indices),
                                                   np.ascontiguousarray(data),
                                                   np.ascontiguousarray(x, data.dtype), y, inner_max)
    return y


def csr_ell_row(N, rowptr, col, val, nc=None):
    """Row-major ELL (par/src/sparse32.js:1840-1880): (nc, indices, data)."""
    rowptr, col, val = _i32(rowptr), _i32(col), np.ascontiguousarray(val)
    if nc is None:
        nc = num_cols(N, rowptr)
    idx = np.zeros(nc * N, np.int32)
    data = np.zeros(nc * N, val.dtype)
    getattr(lib(), "oracle_csr_ell_row_" + _suf(val))(N, nc, rowptr, col, val, idx, data)
    return nc, idx, data


def csr_dia_row(N, rowptr, col, val):
    """Row-major DIA (par/src/sparse32.js:1912-1973): (nd, offset, data)."""
    rowptr, col, val = _i32(rowptr), _i32(col), np.ascontiguousarray(val)
    nd, _ = num_diags(N, rowptr, col)
    off = np.zeros(max(nd, 1), np.int32)
    data = np.zeros(max(nd * N, 1), val.dtype)
    getattr(lib(), "oracle_csr_dia_row_" + _suf(val))(N, nd, rowptr, col, val, off, data)
    return nd, off[:nd], data[:nd * N]


def spmv_dia_row(N, offset, data, x, y, r0=0, r1=None, inner_max=1):
    y = np.array(y, dtype=data.dtype, copy=True)
    r1 = N if r1 is None else r1
    getattr(lib(), "oracle_spmv_dia_row_" + _suf(data))(N, len(offset), _i32(offset), np.ascontiguousarray(data),
                                                       r0, r1, np.ascontiguousarray(x, data.dtype), y, inner_max)
    return y


def spmv_ell_row(N, nc, indices, data, x, y, r0=0, r1=None, gs=False, inner_max=1):
    y = np.array(y, dtype=data.dtype, copy=True)
    r1 = N if r1 is None else r1
    getattr(lib(), "oracle_spmv_ell_row_" + _suf(data))(N, nc, _i32(indices), np.ascontiguousarray(data), r0, r1,
                                                       np.ascontiguousarray(x, data.dtype), y, int(gs), inner_max)
    return y


def spmv_csr_fp64(N, rowptr, col, val, x, y0=None):
    """FP64 oracle: returns (y, sum_j |a_ij x_j|) computed in long double."""
    val = np.ascontiguousarray(val, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros(N, np.float64)
    ab = np.zeros(N, np.float64)
    y0p = None
    if y0 is not None:
        y0 = np.ascontiguousarray(y0, np.float64)
        y0p = y0.ctypes.data
    lib().oracle_spmv_csr_ld(N, _i32(rowptr), _i32(col), val, x, y0p, out, ab.ctypes.data)
    return out, ab


# ------------------------------------------------------------------------ spts
def spts_level(N, rowptr, col, val, b):
    """Level-set order on a (reordered) lower CSR with b in the same order."""
    y = np.zeros(N, val.dtype)
    getattr(lib(), "oracle_spts_level_" + _suf(val))(N, _i32(rowptr), _i32(col),
                                                    np.ascontiguousarray(val),
                                                    np.ascontiguousarray(b, val.dtype), y)
    return y


def spts_syncfree(N, rowptr, col, val, b):
    y = np.zeros(N, val.dtype)
    getattr(lib(), "oracle_spts_syncfree_" + _suf(val))(N, _i32(rowptr), _i32(col),
                                                       np.ascontiguousarray(val),
                                                       np.ascontiguousarray(b, val.dtype), y)
    return y


def spts_csc(N, colptr, row, val, b):
    """seq/src/sparse_opt_32.wat:11-146: lower L in CSC, diagonal first."""
    y = np.zeros(N, val.dtype)
    getattr(lib(), "oracle_spts_csc_" + _suf(val))(N, _i32(colptr), _i32(row), np.ascontiguousarray(val),
                                                  np.ascontiguousarray(b, val.dtype), y)
    return y


def coo_csc(N, row, col, val):
    """seq/src/sparse32.js:1629-1671 -> (colptr, row, val): the counting
    scatter by column plus per-column stable sort by row, i.e. coo_csr with
    the roles of row and column exchanged."""
    return coo_csr(N, col, row, val)


def spts_fp64(N, rowptr, col, val, b):
    y = np.zeros(N, np.float64)
    lib().oracle_spts_ld(N, _i32(rowptr), _i32(col), np.ascontiguousarray(val, np.float64),
                         np.ascontiguousarray(b, np.float64), y)
    return y


def rel_residual(N, rowptr, col, val, x, b):
    return lib().oracle_rel_residual(N, _i32(rowptr), _i32(col),
                                     np.ascontiguousarray(val, np.float64),
                                     np.ascontiguousarray(x, np.float64),
                                     np.ascontiguousarray(b, np.float64))


def fletcher_sum(y):
    y = np.ascontiguousarray(y)
    return getattr(lib(), "oracle_fletcher_" + _suf(y))(len(y), y)


def time_spmv_csr_f64(nrows, rowptr, col, val, x, warm=10, reps=30):
    """Mean seconds per serial CSR SpMV over rows [0, nrows) (C timer)."""
    y = np.zeros(nrows, np.float64)
    t = lib().oracle_time_spmv_csr_f64(nrows, _i32(rowptr), _i32(col), np.ascontiguousarray(val, np.float64),
                                       np.ascontiguousarray(x, np.float64), y, warm, reps)
    return t, y


def static_nnz(nrows, rowptr, workers):
    """Contiguous row ranges of about equal nnz (par/src/sparse32.js:2769-2821)."""
    rs = np.zeros(workers, np.int32)
    re = np.zeros(workers, np.int32)
    lib().oracle_static_nnz(nrows, _i32(rowptr), workers, rs, re)
    return rs, re


def time_spmv_csr_f64_mt(nrows, rowptr, col, val, x, workers, warm=10, reps=30):
    """Mean seconds per CSR SpMV over rows [0, nrows) with `workers`
    persistent threads on static_nnz row ranges (C timer)."""
    y = np.zeros(nrows, np.float64)
    t = lib().oracle_time_spmv_csr_f64_mt(nrows, _i32(rowptr), _i32(col), np.ascontiguousarray(val, np.float64),
                                          np.ascontiguousarray(x, np.float64), y, workers, warm, reps)
    return t, y


def mm_coo(text, dtype=np.float32, lower=False):
    """seq/src/sparse32.js:1889-2172 (32-bit tree) / seq/src/sparse64.js:
    1536-1760 (64-bit): MatrixMarket text -> (info, row, col, val), with
    create_COO_from_MM (lower=False) or create_LCOO_from_MM (lower=True).
    info = dict(nrows, ncols, nentries, nnz, pattern, symmetric).  Raises
    ValueError where the reference throws."""
    if isinstance(text, str):
        text = text.encode()
    dtype = np.dtype(dtype)
    suf = "f32" if dtype == np.float32 else "f64"
    fn = getattr(lib(), "oracle_mm_coo_" + suf)
    info6 = (ctypes.c_int64 * 6)()
    if fn(text, len(text), int(lower), info6, None, None, None):
        raise ValueError("the reference reader throws on this text")
    nnz = int(info6[3])
    row = np.zeros(max(nnz, 1), np.int32)
    col = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1), dtype)
    fn(text, len(text), int(lower), info6, row.ctypes.data, col.ctypes.data, val.ctypes.data)
    info = dict(nrows=int(info6[0]), ncols=int(info6[1]), nentries=int(info6[2]), nnz=nnz,
                pattern=bool(info6[4]), symmetric=bool(info6[5]))
    return info, row[:nnz], col[:nnz], val[:nnz]


# ------------------------------------------------------------------ element-wise
EW_OPS = ["expm1", "log1p", "sin", "tan", "pow", "deg2rad", "rad2deg", "sign", "sign0", "abs", "neg", "sqrt", "ceil",
          "floor", "trunc", "nearest", "multiply"]


def _ewlib(name, suf):
    L = lib()
    fn = getattr(L, "oracle_" + name + "_" + suf)
    return fn


def ew_values(op, param, v):
    """<op>_coo / <op>_ell (seq/src/sparse_opt_32.wat:597-1699, 4977-6347)."""
    v = np.ascontiguousarray(v)
    out = np.empty_like(v)
    fn = _ewlib("ew_values", _suf(v))
    fn.argtypes = [_I, ctypes.c_double, ctypes.c_int64, _VP, _VP]
    fn(EW_OPS.index(op) if isinstance(op, str) else op, float(param), v.size, v.ctypes.data, out.ctypes.data)
    return out


def ew_dia(op, param, N, stride, offset, data, out=None):
    """<op>_dia (seq/src/sparse_opt_32.wat:2137-4976): in-range slots only;
    other slots of out are left as given (zeros by default)."""
    data = np.ascontiguousarray(data)
    out = np.zeros_like(data) if out is None else np.ascontiguousarray(out).copy()
    fn = _ewlib("ew_dia", _suf(data))
    fn.argtypes = [_I, ctypes.c_double, _I, _I, _I, _I32P, _VP, _VP]
    fn(EW_OPS.index(op) if isinstance(op, str) else op, float(param), N, len(offset), stride, _i32(offset),
       data.ctypes.data, out.ctypes.data)
    return out


def count_nonzeros(v):
    v = np.ascontiguousarray(v)
    fn = _ewlib("count_nonzeros", _suf(v))
    fn.argtypes = [ctypes.c_int64, _VP]
    fn.restype = ctypes.c_int64
    return int(fn(v.size, v.ctypes.data))


def copy_nonzeros_coo(row, col, val):
    val = np.ascontiguousarray(val)
    c = count_nonzeros(val)
    orow, ocol, oval = np.zeros(c, np.int32), np.zeros(c, np.int32), np.zeros(c, val.dtype)
    fn = _ewlib("copy_nonzeros_coo", _suf(val))
    fn.argtypes = [ctypes.c_int64, _I32P, _I32P, _VP, _VP, _VP, _VP]
    fn(val.size, _i32(row), _i32(col), val.ctypes.data, orow.ctypes.data, ocol.ctypes.data, oval.ctypes.data)
    return orow, ocol, oval


def copy_nonzeros_csr(N, rowptr, col, val):
    val = np.ascontiguousarray(val)
    c = count_nonzeros(val)
    orp, ocol, oval = np.zeros(N + 1, np.int32), np.zeros(c, np.int32), np.zeros(c, val.dtype)
    fn = _ewlib("copy_nonzeros_csr", _suf(val))
    fn.argtypes = [_I, _I32P, _I32P, _VP, _VP, _VP, _VP]
    fn(N, _i32(rowptr), _i32(col), val.ctypes.data, orp.ctypes.data, ocol.ctypes.data, oval.ctypes.data)
    return orp, ocol, oval


def count_nonzeros_dia(N, stride, offset, data):
    data = np.ascontiguousarray(data)
    out2 = np.zeros(2, np.int64)
    fn = _ewlib("count_nonzeros_dia", _suf(data))
    fn.argtypes = [_I, _I, _I, _I32P, _VP, _VP]
    fn(N, len(offset), stride, _i32(offset), data.ctypes.data, out2.ctypes.data)
    return int(out2[0]), int(out2[1])


def copy_nonzeros_dia(N, stride, offset, data):
    data = np.ascontiguousarray(data)
    cd, _ = count_nonzeros_dia(N, stride, offset, data)
    ooff = np.zeros(cd, np.int32)
    odata = np.zeros(cd * stride, data.dtype)
    fn = _ewlib("copy_nonzeros_dia", _suf(data))
    fn.argtypes = [_I, _I, _I, _I32P, _VP, _VP, _VP, ctypes.c_int64]
    fn(N, len(offset), stride, _i32(offset), data.ctypes.data, ooff.ctypes.data, odata.ctypes.data, odata.size)
    return ooff, odata


def count_nonzeros_ell(N, nc, data):
    data = np.ascontiguousarray(data)
    out2 = np.zeros(2, np.int64)
    fn = _ewlib("count_nonzeros_ell", _suf(data))
    fn.argtypes = [_I, _I, _VP, _VP]
    fn(N, nc, data.ctypes.data, out2.ctypes.data)
    return int(out2[0]), int(out2[1])


def copy_nonzeros_ell(N, nc, data, idx):
    data = np.ascontiguousarray(data)
    cc, _ = count_nonzeros_ell(N, nc, data)
    odata = np.zeros(cc * N, data.dtype)
    oidx = np.zeros(cc * N, np.int32)
    fn = _ewlib("copy_nonzeros_ell", _suf(data))
    fn.argtypes = [_I, _I, _I, _VP, _I32P, _VP, _VP]
    fn(N, nc, cc, data.ctypes.data, _i32(idx), odata.ctypes.data, oidx.ctypes.data)
    return cc, odata, oidx


# ------------------------------------------------------------------ row reordering
def sort_rows_by_nnz(N, rowptr, col, val):
    """par/src/sparse32.js:2148-2218 -> (rowptr, col, val, perm, nnz_row, bins4)."""
    val = np.ascontiguousarray(val)
    nnz = len(val)
    orp, ocol, oval = np.zeros(N + 1, np.int32), np.zeros(nnz, np.int32), np.zeros(nnz, val.dtype)
    perm, nnz_row, bins = np.zeros(N, np.int32), np.zeros(N, np.int32), np.zeros(4, np.int32)
    fn = getattr(lib(), "oracle_sort_rows_by_nnz_" + _suf(val))
    fn.argtypes = [_I, _I32P, _I32P, _VP, _VP, _VP, _VP, _VP, _VP, _VP]
    fn(N, _i32(rowptr), _i32(col), val.ctypes.data, orp.ctypes.data, ocol.ctypes.data, oval.ctypes.data,
       perm.ctypes.data, nnz_row.ctypes.data, bins.ctypes.data)
    return orp, ocol, oval, perm, nnz_row, bins


def sort_y_rows_by_nnz(perm, y):
    """par/src/sparse32.js:2220-2235: out[perm[i]] = y[i]."""
    y = np.ascontiguousarray(y)
    out = np.zeros_like(y)
    fn = getattr(lib(), "oracle_sort_y_rows_" + _suf(y))
    fn.argtypes = [_I, _I32P, _VP, _VP]
    fn(len(perm), _i32(perm), y.ctypes.data, out.ctypes.data)
    return out


def reorder_nn(N, rowptr, col, w=16):
    """par/src/sparse32.js:2501-2622 on a CSR with rows sorted by nnz: the
    greedy window ordering of each bucket (perm[i] = input row at new row i)."""
    perm = np.zeros(N, np.int32)
    fn = lib().oracle_reorder_nn
    fn.argtypes = [_I, _I32P, _I32P, _I, _VP]
    fn(N, _i32(rowptr), _i32(col), w, perm.ctypes.data)
    return perm
