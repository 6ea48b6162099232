"""GPU parity of SpMV (CSR, COO, DIA, ELL) against the CPU oracle.

Tolerance (CSR and COO, whose GPU summation order differs from the
reference's sequential order): for a row with k stored entries both results
lie within gamma_k * (|y0| + sum_j |a_ij x_j|) of the exact value, with
gamma_k = k u / (1 - k u) and u the unit roundoff of the dtype, so
|y_gpu - y_ref| <= 2 gamma_{k+1} (|y0| + sum_j |a_ij x_j|) per row.
DIA and ELL keep the reference's per-row order and its separately rounded
multiply and add: they must be bit-identical.  Every case also reports the
error against the long-double FP64 oracle."""
import numpy as np
import pytest

import oracle
import sciwasm.sparse_amd as sp
from sciwasm.sparse_amd import gen, SableError
from sciwasm.sparse_amd.core import DeviceView

pytestmark = pytest.mark.gpu


def bound(rp, col, val, x, y0, dtype):
    u = np.finfo(dtype).eps / 2
    k = np.diff(rp).astype(np.float64) + 1
    g = k * u / (1 - k * u)
    ab = np.abs(val.astype(np.float64) * x.astype(np.float64)[col])
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    s = np.bincount(rows, weights=ab, minlength=len(rp) - 1) + np.abs(y0.astype(np.float64))
    return 2 * g * s + np.finfo(dtype).tiny


def check_close(y, yref, bnd):
    d = np.abs(y.astype(np.float64) - yref.astype(np.float64))
    bad = np.nonzero(~(d <= bnd))[0]
    assert len(bad) == 0, "rows %s: gpu %s ref %s bound %s" % (bad[:5], y[bad[:5]], yref[bad[:5]], bnd[bad[:5]])


def fp64_report(N, rp, col, val, x, y0, y):
    ye, ab = oracle.spmv_csr_fp64(N, rp, col, val, x, y0)
    err = np.abs(y.astype(np.float64) - ye) / np.maximum(ab + np.abs(np.asarray(y0, np.float64)), 1e-300)
    return float(err.max()) if len(err) else 0.0


CASES = [
    ("laplacian2d_256", lambda dt: gen.laplacian_2d(256, dt)),
    ("random_empty_rows", lambda dt: gen.random_csr(3000, 7, seed=1, dtype=dt, empty_rows=0.2)),
    ("zipf_64k_long_rows", lambda dt: gen.zipf_csr(1 << 16, 1.5, 4096, seed=2, dtype=dt)),
    ("single_row", lambda dt: (np.array([0, 3], np.int32), np.array([0, 0, 0], np.int32),
                               np.array([1.0, 2.0, 3.0], dt))),
]


@pytest.fixture(params=[0, 1, 2, 3], ids=["basic", "staged", "persistent", "streaming"])
def csr_kernel(request):
    from sciwasm.sparse_amd import _lib
    _lib.call("sable_csr_kernel_select", request.param)
    yield request.param
    _lib.call("sable_csr_kernel_select", 0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("name,make", CASES)
@pytest.mark.parametrize("inner_max", [1, 3])
def test_csr_parity(ctx, csr_kernel, dtype, name, make, inner_max):
    rp, col, val = make(dtype)
    N = len(rp) - 1
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, N).astype(dtype)
    y0 = rng.uniform(-1, 1, N).astype(dtype)
    A = sp.csr_from_host(N, rp, col, val)
    xv, yv = sp.vec_from_host(x), sp.vec_from_host(y0)
    sp.spmv(A, xv, yv, inner_max)
    y = yv.array()
    yref = oracle.spmv_csr(N, rp, col, val, x, y0, inner_max)
    b = bound(rp, col, val, x, y0, dtype) * inner_max + bound(rp, col, val, x, yref, dtype) * (inner_max - 1)
    check_close(y, yref, b)
    # the unplanned entry point gives the same bits as the planned one
    yv2 = sp.vec_from_host(y0)
    sp.spmv(A, xv, yv2, inner_max, use_plan=False)
    assert np.array_equal(yv2.array(), y)
    if inner_max == 1:
        e = fp64_report(N, rp, col, val, x, y0, y)
        assert e <= 2 * (np.diff(rp).max() + 2) * np.finfo(dtype).eps, e


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_csr_beta_zero_and_general(ctx, csr_kernel, dtype):
    rp, col, val = gen.random_csr(2000, 9, seed=3, dtype=dtype)
    N = len(rp) - 1
    rng = np.random.default_rng(8)
    x = rng.uniform(-1, 1, N).astype(dtype)
    y0 = rng.uniform(-1, 1, N).astype(dtype)
    A = sp.csr_from_host(N, rp, col, val)
    xv = sp.vec_from_host(x)
    for beta in (0.0, 2.5):
        yv = sp.vec_from_host(y0)
        sp.spmv(A, xv, yv, beta=beta)
        ye, _ = oracle.spmv_csr_fp64(N, rp, col, val, x, beta * y0.astype(np.float64))
        check_close(yv.array(), ye, bound(rp, col, val, x, beta * y0, dtype))


def test_csr_row_slice_like_worker(ctx):
    # par/src/worker32.js:143-155 passes rowptr + 4*start and y + 4*start
    dtype = np.float64
    rp, col, val = gen.random_csr(5000, 6, seed=4, dtype=dtype)
    N = len(rp) - 1
    x = np.random.default_rng(9).uniform(-1, 1, N)
    A = sp.csr_from_host(N, rp, col, val)
    xv, yv = sp.vec_from_host(x), sp.create_vec(N, 0.0, dtype)
    r0, r1 = 1234, 4321
    from sciwasm.sparse_amd import _lib
    _lib.call("sable_spmv_csr", ctx.handle, _lib.F64, r1 - r0, DeviceView(A.row, r0, r1 - r0 + 1).ptr,
              A.col.ptr, A.val.ptr, xv.vec.ptr, DeviceView(yv.vec, r0, r1 - r0).ptr, 1.0, 1)
    y = yv.array()
    yref = oracle.spmv_csr(N, rp, col, val, x, np.zeros(N))
    assert np.all(y[:r0] == 0) and np.all(y[r1:] == 0)
    check_close(y[r0:r1], yref[r0:r1], bound(rp, col, val, x, np.zeros(N), dtype)[r0:r1])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("stride_mode", ["N", "N-min"])
def test_dia_bit_exact(ctx, dtype, stride_mode):
    N = 70000
    rp, col, val = gen.band_csr(N, [-4096, -64, -1, 0, 1, 64, 4096] if stride_mode == "N" else [-300, -2, 3, 77],
                                seed=5, dtype=dtype)
    nd, mn = oracle.num_diags(N, rp, col)
    stride = N if stride_mode == "N" else N - mn
    nd, stride, off, data = oracle.csr_dia(N, rp, col, val, stride=stride)
    rng = np.random.default_rng(10)
    x = rng.uniform(-1, 1, N).astype(dtype)
    y0 = rng.uniform(-1, 1, N).astype(dtype)
    A = sp.dia_from_host(N, stride, off, data, len(col))
    yv = sp.vec_from_host(y0)
    sp.spmv(A, sp.vec_from_host(x), yv, 2)
    assert np.array_equal(yv.array(), oracle.spmv_dia(N, stride, off, data, x, y0, 2))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_ell_bit_exact(ctx, dtype):
    N = 50000
    rp, col, val = gen.random_csr(N, 6, seed=6, dtype=dtype)
    nc, idx, data = oracle.csr_ell(N, rp, col, val)
    rng = np.random.default_rng(11)
    x = rng.uniform(-1, 1, N).astype(dtype)
    y0 = rng.uniform(-1, 1, N).astype(dtype)
    A = sp.ell_from_host(N, nc, idx, data, len(col))
    yv = sp.vec_from_host(y0)
    sp.spmv(A, sp.vec_from_host(x), yv, 2)
    assert np.array_equal(yv.array(), oracle.spmv_ell(N, nc, idx, data, x, y0, 2))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("order", ["sorted", "shuffled"])
def test_coo_parity(ctx, dtype, order):
    rp, col, val = gen.random_csr(20000, 8, seed=12, dtype=dtype, empty_rows=0.1)
    N = len(rp) - 1
    row = gen.csr_to_coo_rows(rp)
    if order == "shuffled":
        p = np.random.default_rng(13).permutation(len(row))
        row, col, val = row[p], col[p], val[p]
    rng = np.random.default_rng(14)
    x = rng.uniform(-1, 1, N).astype(dtype)
    y0 = rng.uniform(-1, 1, N).astype(dtype)
    A = sp.coo_from_host(N, row, col, val)
    yv = sp.vec_from_host(y0)
    sp.spmv(A, sp.vec_from_host(x), yv)
    yref = oracle.spmv_coo(row, col, val, x, y0)
    check_close(yv.array(), yref, bound_coo(N, row, col, val, x, y0, dtype))


def bound_coo(N, row, col, val, x, y0, dtype):
    u = np.finfo(dtype).eps / 2
    k = np.bincount(row, minlength=N).astype(np.float64) + 1
    g = k * u / (1 - k * u)
    s = np.bincount(row, weights=np.abs(val.astype(np.float64) * x.astype(np.float64)[col]), minlength=N)
    return 2 * g * (s + np.abs(y0.astype(np.float64))) + np.finfo(dtype).tiny


def test_coo_empty_and_errors(ctx):
    N = 10
    A = sp.coo_from_host(N, np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0))
    yv = sp.vec_from_host(np.arange(N, dtype=np.float64))
    sp.spmv(A, sp.create_vec(N, 1.0, np.float64), yv)
    assert np.array_equal(yv.array(), np.arange(N))
    B = sp.coo_from_host(N, np.array([0, 11], np.int32), np.array([0, 0], np.int32), np.ones(2))
    with pytest.raises(SableError) as e:
        sp.spmv(B, sp.create_vec(N, 1.0, np.float64), yv)
    assert e.value.name == "SABLE_EMALFORMED"


@pytest.mark.parametrize("lanes,rows", [(2, 1), (8, 4), (8, 16), (64, 2)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("name,make", CASES)
def test_csr_vector_kernel_parity(ctx, lanes, rows, dtype, name, make):
    """The row-vector kernel (G lanes per row, rows longer than G loop) meets
    the CSR bound, y += A.x and y = A.x."""
    rp, col, val = make(dtype)
    N = len(rp) - 1
    rng = np.random.default_rng(41)
    x = rng.uniform(-1, 1, N).astype(dtype)
    y0 = rng.uniform(-1, 1, N).astype(dtype)
    A = sp.csr_from_host(N, rp, col, val)
    pl = sp.CsrPlan(A, {"vector_lanes": lanes, "vector_rows": rows})
    xv = sp.vec_from_host(x)
    for beta in (1.0, 0.0):
        yv = sp.vec_from_host(y0)
        pl.run(xv.vec, yv.vec, beta=beta)
        yb = y0 if beta else np.zeros(N, dtype)
        yref = oracle.spmv_csr(N, rp, col, val, x, yb)
        check_close(yv.array(), yref, bound(rp, col, val, x, yb, dtype))


@pytest.fixture(params=[0, 1], ids=["perpanel", "fused"])
def tiles_mode(request):
    from sciwasm.sparse_amd import _lib
    _lib.call("sable_csr_tiles_select", request.param)
    yield request.param
    _lib.call("sable_csr_tiles_select", 1)


@pytest.mark.parametrize("tile_rows,panel_cols", [(64, 128), (192, 256), (256, 1024), (16384, 8192), (20480, 8192),
                                                 (0, 0)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("name,make", CASES)
def test_csr_tiles_parity(ctx, tiles_mode, tile_rows, panel_cols, dtype, name, make):
    """Column-sorted tiles: y = beta*y + A.x with beta 1, 0 and 0.5, and
    inner_max = 2, within the CSR bound (small tiles and panels exercise
    many tiles, partial last tiles and panels, and empty tiles)."""
    rp, col, val = make(dtype)
    N = len(rp) - 1
    rng = np.random.default_rng(51)
    x = rng.uniform(-1, 1, N).astype(dtype)
    y0 = rng.uniform(-1, 1, N).astype(dtype)
    A = sp.csr_from_host(N, rp, col, val)
    pl = sp.CsrTiles(A, tile_rows, panel_cols)
    info = pl.info()
    trr = tile_rows or 16384
    pcc = panel_cols or 8192
    assert info["row_tiles"] == -(-N // trr) and info["npanels"] == -(-N // pcc)
    xv = sp.vec_from_host(x)
    for beta in (1.0, 0.0, 0.5):
        yv = sp.vec_from_host(y0)
        pl.run(xv.vec, yv.vec, beta=beta)
        yb = (beta * y0.astype(np.float64)).astype(dtype)
        yref = oracle.spmv_csr(N, rp, col, val, x, yb)
        check_close(yv.array(), yref, bound(rp, col, val, x, yb, dtype) + np.finfo(dtype).eps * np.abs(y0))
    yv = sp.vec_from_host(y0)
    pl.run(xv.vec, yv.vec, beta=1.0, inner_max=2)
    y1 = oracle.spmv_csr(N, rp, col, val, x, y0, 1)
    y2 = oracle.spmv_csr(N, rp, col, val, x, y0, 2)
    check_close(yv.array(), y2, bound(rp, col, val, x, y0, dtype) + bound(rp, col, val, x, y1, dtype))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("name,make", CASES)
def test_csr_tiles_from_sorted_coo(ctx, dtype, name, make):
    """A tile plan built from a row-sorted COO (its plan's row heads with
    its col and val), or from the same entries shuffled (sorted first),
    gives the CSR product within the CSR bound; an unsorted COO's plan has
    no row heads."""
    rp, col, val = make(dtype)
    N = len(rp) - 1
    row = np.repeat(np.arange(N), np.diff(rp)).astype(np.int32)
    rng = np.random.default_rng(53)
    x = rng.uniform(-1, 1, N).astype(dtype)
    C = sp.coo_from_host(N, row, col, val)
    pl = sp.CsrTiles.from_coo(C, 256, 1024)
    yv = sp.create_vec(N, 0.0, dtype)
    pl.run(sp.vec_from_host(x).vec, yv.vec, beta=0.0)
    z = np.zeros(N, dtype)
    check_close(yv.array(), oracle.spmv_csr(N, rp, col, val, x, z), bound(rp, col, val, x, z, dtype))
    if len(row) > 1 and row[0] != row[-1]:
        p = np.random.default_rng(54).permutation(len(row))
        U = sp.coo_from_host(N, row[p], col[p], val[p])
        with pytest.raises(SableError) as e:
            sp.CooPlan(U).row_ptr()
        assert e.value.name == "SABLE_EINVAL"
        pu = sp.CsrTiles.from_coo(U, 256, 1024)
        yu = sp.create_vec(N, 0.0, dtype)
        pu.run(sp.vec_from_host(x).vec, yu.vec, beta=0.0)
        check_close(yu.array(), oracle.spmv_csr(N, rp, col, val, x, z), bound(rp, col, val, x, z, dtype))


def test_csr_tiles_rejects_bad_shapes(ctx):
    rp, col, val = gen.laplacian_2d(16)
    A = sp.csr_from_host(256, rp, col, val)
    for tr, pc in ((100, 128), (1 << 17, 1 << 16), (1 << 15, 1 << 10)):
        with pytest.raises(SableError) as e:
            sp.CsrTiles(A, tr, pc)
        assert e.value.name == "SABLE_EINVAL", (tr, pc)


@pytest.mark.parametrize("order", ["sorted", "shuffled"])
def test_coo_plan_matches_unplanned(ctx, order):
    """The planned product (rows analysed once) equals the unplanned entry
    point: bit for bit on row-sorted input (both run the same segmented
    kernel), within the COO bound on unsorted input (atomics)."""
    rp, col, val = gen.random_csr(9000, 6, seed=31, empty_rows=0.1)
    N = len(rp) - 1
    row = gen.csr_to_coo_rows(rp)
    if order == "shuffled":
        p = np.random.default_rng(32).permutation(len(row))
        row, col, val = row[p], col[p], val[p]
    x = np.random.default_rng(33).uniform(-1, 1, N)
    y0 = np.random.default_rng(34).uniform(-1, 1, N)
    A = sp.coo_from_host(N, row, col, val)
    assert A.plan().info()["sorted"] == (order == "sorted")
    ya, yb = sp.vec_from_host(y0), sp.vec_from_host(y0)
    sp.spmv(A, sp.vec_from_host(x), ya, inner_max=2)
    sp.spmv(A, sp.vec_from_host(x), yb, inner_max=2, use_plan=False)
    if order == "sorted":
        assert np.array_equal(ya.array(), yb.array())
    yref = oracle.spmv_coo(row, col, val, x, y0, 2)
    y1 = oracle.spmv_coo(row, col, val, x, y0, 1)
    b = bound_coo(N, row, col, val, x, y0, np.float64) + bound_coo(N, row, col, val, x, y1, np.float64)
    check_close(ya.array(), yref, b)


def test_csr_malformed_row_ptr(ctx):
    rp = np.array([0, 3, 2, 4], np.int32)
    A = sp.csr_from_host(3, rp, np.zeros(4, np.int32), np.ones(4))
    with pytest.raises(SableError) as e:
        A.plan()
    assert e.value.name == "SABLE_EMALFORMED"


def test_ell_zero_width_leaves_y(ctx):
    A = sp.ell_from_host(5, 0, np.zeros(0, np.int32), np.zeros(0), 0)
    yv = sp.vec_from_host(np.arange(5, dtype=np.float64))
    sp.spmv(A, sp.create_vec(5, 1.0, np.float64), yv)
    assert np.array_equal(yv.array(), np.arange(5))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("panel_cols", [1000, 4096, 0])
@pytest.mark.parametrize("beta", [0.0, 1.0])
def test_csr_panels_parity(ctx, csr_kernel, dtype, panel_cols, beta):
    rp, col, val = gen.zipf_csr(1 << 16, 1.5, 4096, seed=31, dtype=dtype)
    N = len(rp) - 1
    rng = np.random.default_rng(12)
    x = rng.uniform(-1, 1, N).astype(dtype)
    y0 = rng.uniform(-1, 1, N).astype(dtype)
    A = sp.csr_from_host(N, rp, col, val)
    P = sp.CsrPanels(A, panel_cols=panel_cols)
    info = P.info()
    assert info["npanels"] == (1 if panel_cols == 0 else -(-N // panel_cols))
    yv = sp.vec_from_host(y0)
    P.run(sp.vec_from_host(x).vec, yv.vec, beta=beta, inner_max=2 if beta else 1)
    y = yv.array()
    if beta:
        yref = oracle.spmv_csr(N, rp, col, val, x, y0, 2)
        b = bound(rp, col, val, x, y0, dtype) * 2 + bound(rp, col, val, x, yref, dtype)
    else:
        yref = oracle.spmv_csr(N, rp, col, val, x, np.zeros(N, dtype))
        b = bound(rp, col, val, x, np.zeros(N, dtype), dtype)
    check_close(y, yref, b)
