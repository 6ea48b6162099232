"""Element-wise ops and eliminate_zeros on the GPU against the oracle.

Bit-exact: every op that is a WebAssembly instruction (abs, neg, sqrt,
ceil, floor, trunc, nearest, deg2rad, rad2deg, sign, multiply) and all
compactions.  expm1, log1p, sin, tan and pow call JavaScript's Math in the
reference, whose results are not specified to the last bit: the device's
double-precision library and the C library may differ by one unit in the
double result, so FP64 values agree within 2 ulp and FP32 values (rounded
from the double) within 1 ulp, with NaN exactly where the oracle has NaN.
"""
import math

import numpy as np
import pytest

import oracle
import sciwasm.sparse_amd as sp
from sciwasm.sparse_amd._lib import call

from test_ewise import sample

pytestmark = pytest.mark.gpu

DT = [np.float32, np.float64]
EXACT = ["abs", "neg", "sqrt", "ceil", "floor", "trunc", "nearest", "deg2rad", "rad2deg", "sign"]
TRANS = ["expm1", "log1p", "sin", "tan"]


def ulps(a, b):
    a, b = np.asarray(a), np.asarray(b)
    it = np.int32 if a.dtype == np.float32 else np.int64
    ai, bi = a.view(it).astype(np.int64), b.view(it).astype(np.int64)
    # map the sign-magnitude order onto integers
    ai = np.where(ai < 0, np.iinfo(it).min - ai, ai)
    bi = np.where(bi < 0, np.iinfo(it).min - bi, bi)
    return np.abs(ai - bi)


def check(got, want, op):
    got, want = np.asarray(got), np.asarray(want)
    assert got.dtype == want.dtype
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan), op
    g, w = got[~nan], want[~nan]
    if op in TRANS or op == "pow":
        tol = 1 if got.dtype == np.float32 else 2
        inf = np.isinf(w)
        assert np.array_equal(g[inf], w[inf]), op
        assert np.all(ulps(g[~inf], w[~inf]) <= tol), (op, int(ulps(g[~inf], w[~inf]).max()))
    else:
        assert np.array_equal(g.view(np.uint8), w.view(np.uint8)), op


def param(op):
    return math.pi if op in ("deg2rad", "rad2deg") else 0.0


def coo_of(dtype, seed=0):
    v = sample(dtype, 3000, seed)
    rng = np.random.default_rng(seed)
    N = 400
    row = np.sort(rng.integers(0, N, len(v))).astype(np.int32)
    col = rng.integers(0, N, len(v)).astype(np.int32)
    return N, row, col, v


@pytest.mark.parametrize("dtype", DT, ids=["f32", "f64"])
@pytest.mark.parametrize("op", EXACT + TRANS + ["pow"])
def test_ew_coo_csr(ctx, dtype, op):
    N, row, col, v = coo_of(dtype)
    A = sp.coo_from_host(N, row, col, v, ctx)
    p = 1.7 if op == "pow" else 0.0
    B = sp.pow(A, p) if op == "pow" else getattr(sp, op)(A)
    want = oracle.ew_values(op, p if op == "pow" else param(op), v)
    check(B.val.numpy(), want, op)
    assert np.array_equal(B.row.numpy(), row) and np.array_equal(B.col.numpy(), col)
    assert np.array_equal(A.val.numpy().view(np.uint8), v.view(np.uint8))  # input untouched
    # CSR values take the same path; in place on the object
    rp = np.searchsorted(row, np.arange(N + 1)).astype(np.int32)
    C = sp.csr_from_host(N, rp, col, v, ctx)
    (C.pow(p) if op == "pow" else getattr(C, op)())
    check(C.val.numpy(), want, op)


@pytest.mark.parametrize("dtype", DT, ids=["f32", "f64"])
@pytest.mark.parametrize("op", EXACT + ["expm1", "pow"])
def test_ew_dia(ctx, dtype, op):
    N = 300
    off = np.array([-7, -1, 0, 3, 250], np.int32)
    stride = N
    data = sample(dtype, len(off) * stride, 3)
    D = sp.dia_from_host(N, stride, off, data, 1000, ctx)
    p = -0.5 if op == "pow" else 0.0
    B = sp.pow(D, p) if op == "pow" else getattr(sp, op)(D)
    pa = p if op == "pow" else param(op)
    opname = "sign0" if op == "sign" else op
    check(B.data.numpy(), oracle.ew_dia(opname, pa, N, stride, off, data), op)  # padding stays zero
    # in place: padding slots keep their values
    (D.pow(p) if op == "pow" else getattr(D, op)())
    check(D.data.numpy(), oracle.ew_dia(opname, pa, N, stride, off, data, out=data), op)


@pytest.mark.parametrize("dtype", DT, ids=["f32", "f64"])
@pytest.mark.parametrize("op", EXACT + ["sin", "pow"])
def test_ew_ell(ctx, dtype, op):
    N, nc = 500, 4
    data = sample(dtype, nc * N, 4)
    idx = np.random.default_rng(4).integers(0, N, nc * N).astype(np.int32)
    E = sp.ell_from_host(N, nc, idx, data, nc * N, ctx)
    p = 0.0 if op == "pow" else 0.0  # pow(0, 0) = 1: padding becomes 1, as in the reference
    B = sp.pow(E, p) if op == "pow" else getattr(sp, op)(E)
    opname = "sign0" if op == "sign" else op
    check(B.data.numpy(), oracle.ew_values(opname, p if op == "pow" else param(op), data), op)
    assert np.array_equal(B.indices.numpy(), idx)


def test_multiply_scalar_coo(ctx):
    N, row, col, v = coo_of(np.float32, 5)
    A = sp.coo_from_host(N, row, col, v, ctx)
    A.multiply(-3.25)
    check(A.val.numpy(), oracle.ew_values("multiply", -3.25, v), "multiply")


@pytest.mark.parametrize("dtype", DT, ids=["f32", "f64"])
def test_eliminate_zeros_coo_csr(ctx, dtype):
    N, row, col, v = coo_of(dtype, 6)
    keep_n = oracle.count_nonzeros(v)
    A = sp.coo_from_host(N, row, col, v, ctx)
    assert sp.count_nonzeros(A) == keep_n
    B = sp.eliminate_zeros(A)
    r, c, w = oracle.copy_nonzeros_coo(row, col, v)
    assert B.nnz == keep_n
    assert np.array_equal(B.row.numpy(), r) and np.array_equal(B.col.numpy(), c)
    assert np.array_equal(B.val.numpy().view(np.uint8), w.view(np.uint8))
    rp = np.searchsorted(row, np.arange(N + 1)).astype(np.int32)
    C = sp.csr_from_host(N, rp, col, v, ctx)
    D = sp.eliminate_zeros(C)
    orp, oc, ov = oracle.copy_nonzeros_csr(N, rp, col, v)
    assert np.array_equal(D.row.numpy(), orp) and np.array_equal(D.col.numpy(), oc)
    assert np.array_equal(D.val.numpy().view(np.uint8), ov.view(np.uint8))
    # in place, and a matrix without zeros is left as it is
    C.eliminate_zeros()
    assert C.nnz == keep_n and np.array_equal(C.row.numpy(), orp)
    E = sp.coo_from_host(N, r, c, w, ctx)
    before = E.val.ptr
    E.eliminate_zeros()
    assert E.val.ptr == before and E.nnz == keep_n


@pytest.mark.parametrize("dtype", DT, ids=["f32", "f64"])
def test_eliminate_zeros_dia_ell(ctx, dtype):
    N = 700
    rng = np.random.default_rng(8)
    off = np.array([-5, -2, 0, 1, 9, 600], np.int32)
    data = rng.standard_normal(len(off) * N).astype(dtype)
    data[rng.random(data.size) < 0.35] = 0
    data[3 * N:4 * N] = 0  # the k=1 diagonal is all zero: dropped
    D = sp.dia_from_host(N, N, off, data, int(data.size), ctx)
    cd, cnt = oracle.count_nonzeros_dia(N, N, off, data)
    assert sp.count_nonzeros(D) == (cd, cnt)
    E = sp.eliminate_zeros(D)
    ooff, odata = oracle.copy_nonzeros_dia(N, N, off, data)
    assert (E.ndiags, E.nnz) == (cd, cnt)
    assert np.array_equal(E.offset.numpy(), ooff)
    assert np.array_equal(E.data.numpy().view(np.uint8), odata.view(np.uint8))
    # ELL
    nc = 6
    ed = rng.standard_normal(nc * N).astype(dtype)
    ed[rng.random(ed.size) < 0.5] = 0
    ei = rng.integers(0, N, nc * N).astype(np.int32)
    L = sp.ell_from_host(N, nc, ei, ed, nc * N, ctx)
    cc, odat, oidx = oracle.copy_nonzeros_ell(N, nc, ed, ei)
    assert sp.count_nonzeros(L) == oracle.count_nonzeros_ell(N, nc, ed)
    M = sp.eliminate_zeros(L)
    assert M.ncols == cc
    assert np.array_equal(M.data.numpy().view(np.uint8), odat.view(np.uint8))
    assert np.array_equal(M.indices.numpy(), oidx)
    L.eliminate_zeros()
    assert L.ncols == cc and np.array_equal(L.indices.numpy(), oidx)


def test_self_transpose_coo_sorts(ctx):
    rng = np.random.default_rng(9)
    N = 200
    key = rng.permutation(N * N)[:3000]
    row, col = (key // N).astype(np.int32), (key % N).astype(np.int32)
    val = rng.standard_normal(len(key))
    A = sp.coo_from_host(N, row, col, val, ctx)
    A.transpose()
    # the reference swaps row and col, then quick_sort_COO by (row, col)
    order = np.lexsort((row, col))
    assert np.array_equal(A.row.numpy(), col[order]) and np.array_equal(A.col.numpy(), row[order])
    assert np.array_equal(A.val.numpy(), val[order])


def test_ew_large_bandwidth_path(ctx):
    """A 64M-value array through the grid-stride kernels, in place and out."""
    n = 1 << 26
    v = np.random.default_rng(10).standard_normal(n).astype(np.float32)
    a = sp.DeviceArray.from_host(v, ctx=ctx)
    b = sp.DeviceArray(n, np.float32, ctx)
    call("sable_ew_values", ctx.handle, 0, 10, 0.0, n, a.ptr, b.ptr)  # neg, out of place
    call("sable_ew_values", ctx.handle, 0, 9, 0.0, n, a.ptr, a.ptr)   # abs, in place
    assert np.array_equal(b.numpy(), -v) and np.array_equal(a.numpy(), np.abs(v))
