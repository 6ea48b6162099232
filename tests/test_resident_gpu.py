"""GPU parity of the entry points behind the addon's resident matrices and
the parallel tree's slices, against the CPU oracle:

* sable_csr_auto_* (autotuned schedule): every schedule it may pick meets
  the CSR bound |y - y_ref| <= 2 gamma_{k+1} (|y0| + sum |a x|) per row, for
  full matrices and row slices (row_ptr[0] != 0, columns beyond the slice's
  rows); a second matrix of the same class reuses the cached choice; the
  deterministic flag never picks tiles;
* sable_spmv_csr (plan built on the device, no synchronisation), including
  a matrix whose default plan needs more blocks than its capacity, and a
  decreasing row_ptr that leaves y untouched;
* row ranges of column-major DIA/ELL and the row-major DIA/ELL products of
  the parallel tree (plain and gs order): bit-identical to the oracle, rows
  outside the range untouched;
* the row-major conversions: bit-identical;
* the staged pageable copies and sable_max_i32: exact."""
import ctypes

import numpy as np
import pytest

import oracle
import sciwasm.sparse_amd as sp
from sciwasm.sparse_amd import gen, _lib
from sciwasm.sparse_amd._lib import call
from sciwasm.sparse_amd.core import DeviceArray, dtype_code

pytestmark = pytest.mark.gpu


def bound(rp, col, val, x, y0, dtype):
    u = np.finfo(dtype).eps / 2
    k = np.diff(rp).astype(np.float64) + 1
    g = k * u / (1 - k * u)
    ab = np.abs(val.astype(np.float64) * x.astype(np.float64)[col])
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    s = np.bincount(rows, weights=ab, minlength=len(rp) - 1) + np.abs(y0.astype(np.float64))
    return 2 * g * s + np.finfo(dtype).tiny


def check_close(y, yref, bnd, what=""):
    d = np.abs(y.astype(np.float64) - yref.astype(np.float64))
    bad = np.nonzero(~(d <= bnd))[0]
    assert len(bad) == 0, (what, bad[:5], y[bad[:5]], yref[bad[:5]])


MATS = [
    ("laplacian2d_128", lambda dt: gen.laplacian_2d(128, dt)),
    ("zipf_64k", lambda dt: gen.zipf_csr(1 << 16, 1.5, 4096, seed=3, dtype=dt)),
    ("random_empty_rows", lambda dt: gen.random_csr(5000, 9, seed=4, dtype=dt, empty_rows=0.2)),
    ("band_big_x", lambda dt: gen.band_csr(1 << 20, (0, 1, -1, 64, -64, 4096, -4096), seed=5, dtype=dt)),
]


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("name,make", MATS, ids=[m[0] for m in MATS])
@pytest.mark.parametrize("det", [False, True], ids=["any", "deterministic"])
def test_auto_plan_meets_bound(ctx, dtype, name, make, det):
    print("auto", name, dtype.__name__, det, flush=True)
    rp, col, val = make(dtype)
    N = len(rp) - 1
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, N).astype(dtype)
    y0 = rng.uniform(-1, 1, N).astype(dtype)
    A = sp.csr_from_host(N, rp, col, val)
    p = sp.CsrAuto(A, deterministic=det)
    info = p.info()
    assert info["ncols"] == int(col.max()) + 1 and info["nnz"] == len(col)
    if det:
        assert info["schedule"] != "tiles"
    xd, yd = DeviceArray.from_host(x), DeviceArray.from_host(y0)
    p.run(xd, yd, beta=1.0, inner_max=2)
    yref = oracle.spmv_csr(N, rp, col, val, x, y0, 2)
    b1 = bound(rp, col, val, x, y0, dtype)
    check_close(yd.numpy(), yref, b1 + bound(rp, col, val, x, yref, dtype), name)
    # beta = 0: the plain product
    p.run(xd, yd, beta=0.0)
    check_close(yd.numpy(), oracle.spmv_csr(N, rp, col, val, x, np.zeros(N, dtype), 1),
                bound(rp, col, val, x, np.zeros(N, dtype), dtype), name)
    # the same class again: the cached choice, not timed again
    q = sp.CsrAuto(A, deterministic=det)
    assert q.info()["from_cache"] and q.info()["schedule"] == info["schedule"]


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_auto_plan_row_slice(ctx, dtype):
    """A worker's slice: row_ptr starts inside the matrix, columns range over
    all N; every schedule is timed and the chosen one meets the bound."""
    rp, col, val = gen.zipf_csr(1 << 15, 1.5, 2048, seed=6, dtype=dtype)
    N = len(rp) - 1
    x = np.random.default_rng(2).uniform(-1, 1, N).astype(dtype)
    r0, r1 = 5000, 21000
    A = sp.csr_from_host(N, rp, col, val)
    from sciwasm.sparse_amd.core import sparse_CSR_t
    sub = DeviceArray.from_host(rp[r0:r1 + 1])
    S = sparse_CSR_t(sub, A.col, A.val, r1 - r0, int(rp[r1] - rp[r0]))
    sp.CsrAuto.clear_cache()
    p = sp.CsrAuto(S)
    xd = DeviceArray.from_host(x)
    yd = DeviceArray.zeros(r1 - r0, dtype)
    p.run(xd, yd, beta=0.0)
    srp = (rp[r0:r1 + 1] - rp[r0]).astype(np.int32)
    scol, sval = col[rp[r0]:rp[r1]], val[rp[r0]:rp[r1]]
    z = np.zeros(r1 - r0, dtype)
    check_close(yd.numpy(), oracle.spmv_csr(r1 - r0, srp, scol, sval, x, z, 1), bound(srp, scol, sval, x, z, dtype))


def test_auto_plan_tiles_on_a_slice(ctx):
    """The tile schedule of a row slice (rebased row_ptr, rectangular x)."""
    dtype = np.float64
    rp, col, val = gen.zipf_csr(1 << 17, 1.5, 4096, seed=7, dtype=dtype)
    N = len(rp) - 1
    r0, r1 = 30000, 90000
    from sciwasm.sparse_amd.core import sparse_CSR_t
    A = sp.csr_from_host(N, rp, col, val)
    sub = DeviceArray.from_host(rp[r0:r1 + 1])
    S = sparse_CSR_t(sub, A.col, A.val, r1 - r0, int(rp[r1] - rp[r0]))
    h = ctypes.c_void_p()
    srp = (rp[r0:r1 + 1] - rp[r0]).astype(np.int32)
    scol, sval = col[rp[r0]:rp[r1]], val[rp[r0]:rp[r1]]
    # the rectangular tile plan directly, then the same through the auto plan
    d_rp = DeviceArray.from_host(srp)
    d_c = DeviceArray.from_host(scol)
    d_v = DeviceArray.from_host(sval)
    call("sable_csr_tiles_create_rect", ctx.handle, 1, r1 - r0, N, d_rp.ptr, d_c.ptr, d_v.ptr, 4096, 1 << 17,
         ctypes.byref(h))
    x = np.random.default_rng(3).uniform(-1, 1, N)
    xd, yd = DeviceArray.from_host(x), DeviceArray.zeros(r1 - r0, dtype)
    call("sable_spmv_csr_tiles", ctx.handle, h, xd.ptr, yd.ptr, 0.0, 1)
    z = np.zeros(r1 - r0)
    yref = oracle.spmv_csr(r1 - r0, srp, scol, sval, x, z, 1)
    check_close(yd.numpy(), yref, bound(srp, scol, sval, x, z, dtype))
    _lib.load().sable_csr_tiles_destroy(h)
    p = sp.CsrAuto(S)
    yd.fill(0)
    p.run(xd, yd, beta=0.0)
    check_close(yd.numpy(), yref, bound(srp, scol, sval, x, z, dtype))


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_plan_free_csr_heavy_rows_and_malformed(ctx, dtype):
    """sable_spmv_csr builds its plan on the device: 8 rows of 250 000
    entries need more blocks than its capacity at the default weight, so
    its blocks get heavier; a decreasing row_ptr leaves y unchanged."""
    rng = np.random.default_rng(8)
    N, k = 8, 250000
    rp = (np.arange(N + 1) * k).astype(np.int32)
    col = rng.integers(0, N, N * k).astype(np.int32)
    val = rng.uniform(-1, 1, N * k).astype(dtype)
    x = rng.uniform(-1, 1, N).astype(dtype)
    A = sp.csr_from_host(N, rp, col, val)
    xd, yd = DeviceArray.from_host(x), DeviceArray.zeros(N, dtype)
    call("sable_spmv_csr", ctx.handle, dtype_code(dtype), N, A.row.ptr, A.col.ptr, A.val.ptr, xd.ptr, yd.ptr, 0.0, 1)
    z = np.zeros(N, dtype)
    check_close(yd.numpy(), oracle.spmv_csr(N, rp, col, val, x, z, 1), bound(rp, col, val, x, z, dtype))
    bad = rp.copy()
    bad[3] = bad[5] + 7
    B = DeviceArray.from_host(bad)
    y0 = rng.uniform(-1, 1, N).astype(dtype)
    yd.copy_from(y0)
    call("sable_spmv_csr", ctx.handle, dtype_code(dtype), N, B.ptr, A.col.ptr, A.val.ptr, xd.ptr, yd.ptr, 1.0, 1)
    assert np.array_equal(yd.numpy(), y0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_dia_ell_row_ranges_and_row_major(ctx, dtype):
    rng = np.random.default_rng(9)
    rp, col, _ = gen.laplacian_2d(60)
    N = len(rp) - 1
    val = rng.uniform(-1, 1, len(col)).astype(dtype)
    x = rng.uniform(-1, 1, N).astype(dtype)
    y0 = rng.uniform(-1, 1, N).astype(dtype)
    dt = dtype_code(dtype)
    r0, r1 = 517, 2901
    xd = DeviceArray.from_host(x)

    def run(fn, *args):
        yd = DeviceArray.from_host(y0)
        call(fn, ctx.handle, dt, *[a(yd) if callable(a) else a for a in args])
        return yd.numpy()

    def ranged(yfull):
        out = y0.copy()
        out[r0:r1] = yfull[r0:r1]
        return out

    nd, stride, off, dd = oracle.csr_dia(N, rp, col, val)
    doff, ddat = DeviceArray.from_host(off), DeviceArray.from_host(dd)
    y = run("sable_spmv_dia_range", N, nd, stride, doff.ptr, ddat.ptr, r0, r1, xd.ptr, lambda v: v.ptr, 1.0, 3)
    assert np.array_equal(y.view(np.uint8), ranged(oracle.spmv_dia(N, stride, off, dd, x, y0, 3)).view(np.uint8))
    nc, ci, cd = oracle.csr_ell(N, rp, col, val)
    dci, dcd = DeviceArray.from_host(ci), DeviceArray.from_host(cd)
    y = run("sable_spmv_ell_range", N, nc, dci.ptr, dcd.ptr, r0, r1, xd.ptr, lambda v: v.ptr, 1.0, 2)
    assert np.array_equal(y.view(np.uint8), ranged(oracle.spmv_ell(N, nc, ci, cd, x, y0, 2)).view(np.uint8))

    # row-major conversions on the device
    A = sp.csr_from_host(N, rp, col, val)
    _, ri, rd = oracle.csr_ell_row(N, rp, col, val)
    oi, od = DeviceArray(nc * N, np.int32), DeviceArray(nc * N, dtype)
    call("sable_csr_ell_rowmajor", ctx.handle, dt, N, nc, A.row.ptr, A.col.ptr, A.val.ptr, oi.ptr, od.ptr)
    assert np.array_equal(oi.numpy(), ri) and np.array_equal(od.numpy(), rd)
    _, roff, rdd = oracle.csr_dia_row(N, rp, col, val)
    oo, odd = DeviceArray(nd, np.int32), DeviceArray(nd * N, dtype)
    call("sable_csr_dia_rowmajor", ctx.handle, dt, N, nd, A.row.ptr, A.col.ptr, A.val.ptr, oo.ptr, odd.ptr)
    assert np.array_equal(oo.numpy(), roff) and np.array_equal(odd.numpy(), rdd)

    # row-major products: plain and gs order, with a negative index row
    ri2 = ri.copy()
    ri2[7 * nc + 2] = -1
    dri, drd = DeviceArray.from_host(ri2), DeviceArray.from_host(rd)
    for order in (0, 1):
        y = run("sable_spmv_ell_rowmajor", N, nc, dri.ptr, drd.ptr, r0, r1, xd.ptr, lambda v: v.ptr, order, 2)
        yref = oracle.spmv_ell_row(N, nc, ri2, rd, x, y0, r0=r0, r1=r1, gs=bool(order), inner_max=2)
        assert np.array_equal(y.view(np.uint8), yref.view(np.uint8)), order
    # the negative index sits outside the range: a full sweep meets it
    y = run("sable_spmv_ell_rowmajor", N, nc, dri.ptr, drd.ptr, 0, N, xd.ptr, lambda v: v.ptr, 0, 1)
    assert np.array_equal(y, oracle.spmv_ell_row(N, nc, ri2, rd, x, y0))
    droff, drdd = DeviceArray.from_host(roff), DeviceArray.from_host(rdd)
    y = run("sable_spmv_dia_rowmajor", N, nd, droff.ptr, drdd.ptr, r0, r1, xd.ptr, lambda v: v.ptr, 3)
    yref = oracle.spmv_dia_row(N, roff, rdd, x, y0, r0=r0, r1=r1, inner_max=3)
    assert np.array_equal(y.view(np.uint8), yref.view(np.uint8))


def test_staged_copies_and_max(ctx):
    rng = np.random.default_rng(10)
    for n in (1, 4095, (4 << 20) - 1, 4 << 20, (9 << 20) + 3):
        a = rng.integers(0, 256, n).astype(np.uint8)
        d = DeviceArray(n, np.uint8)
        call("sable_memcpy_h2d_staged", ctx.handle, d.ptr, a.ctypes.data, n)
        b = np.zeros(n, np.uint8)
        call("sable_memcpy_d2h_staged", ctx.handle, b.ctypes.data, d.ptr, n)
        assert np.array_equal(a, b), n
    v = rng.integers(-1000, 1 << 30, 3_000_001).astype(np.int32)
    dv = DeviceArray.from_host(v)
    m = ctypes.c_int32()
    call("sable_max_i32", ctx.handle, dv.ptr, len(v), ctypes.byref(m))
    assert m.value == v.max()
    call("sable_max_i32", ctx.handle, dv.ptr, 0, ctypes.byref(m))
    assert m.value == -1
