"""GPU conversions and transposes: bit-exact against the oracle (and, at
sizes where the oracle's per-row bubble sort is too slow, against the stable
(row, col, input position) sort that test_oracle.py pins it to)."""
import ctypes

import numpy as np
import pytest

import oracle
import sciwasm.sparse_amd as sp
from sciwasm.sparse_amd import gen, _lib, DeviceArray, SableError

pytestmark = pytest.mark.gpu


def csr_host(A):
    return A.row.numpy(), A.col.numpy(), A.val.numpy()


def shuffled_coo(N, nnz, seed, dtype, dup=True):
    rng = np.random.default_rng(seed)
    row = rng.integers(0, N, nnz).astype(np.int32)
    col = rng.integers(0, N, nnz).astype(np.int32)
    val = rng.standard_normal(nnz).astype(dtype)
    return row, col, val


# -------------------------------------------------------------- primitives
@pytest.mark.parametrize("n", [1, 7, 4095, 4096, 4097, 1 << 20, 10_000_019])
def test_scan(ctx, n):
    a = np.random.default_rng(n).integers(0, 50, n).astype(np.int32)
    d = DeviceArray.from_host(a)
    out = DeviceArray(n + 1, np.int32)
    _lib.call("sable_scan_exclusive_i32", ctx.handle, d.ptr, out.ptr, n)
    ref = np.zeros(n + 1, np.int64)
    ref[1:] = np.cumsum(a)
    assert np.array_equal(out.numpy(), ref.astype(np.int32))


@pytest.mark.parametrize("n,bits", [(2, 8), (2047, 12), (2049, 20), (300_001, 44), (3_000_000, 22)])
def test_sort_pairs_stable(ctx, n, bits):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1 << min(bits, 16), n).astype(np.uint64) << np.uint64(max(bits - 16, 0))
    vals = np.arange(n, dtype=np.uint32)
    dk, dv = DeviceArray.from_host(keys), DeviceArray.from_host(vals)
    _lib.call("sable_sort_pairs_u64", ctx.handle, dk.ptr, dv.ptr, n, bits)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(dk.numpy(), keys[order])
    assert np.array_equal(dv.numpy(), vals[order])


@pytest.mark.parametrize("n,bits", [(5000, 12), (70_001, 27), (9000, 8)])
def test_sort_pairs_carries_bits_above_end_bit(ctx, n, bits):
    """Only key bits [0, bits) order the keys; the bits above ride along."""
    rng = np.random.default_rng(bits)
    keys = rng.integers(0, 1 << 40, n).astype(np.uint64)
    vals = np.arange(n, dtype=np.uint32)
    dk, dv = DeviceArray.from_host(keys), DeviceArray.from_host(vals)
    _lib.call("sable_sort_pairs_u64", ctx.handle, dk.ptr, dv.ptr, n, bits)
    order = np.argsort(keys & np.uint64((1 << bits) - 1), kind="stable")
    assert np.array_equal(dk.numpy(), keys[order])
    assert np.array_equal(dv.numpy(), vals[order])


# ----------------------------------------------------------------- coo_csr
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("N,nnz", [(1, 1), (37, 400), (1000, 30000)])
def test_coo_csr_bit_exact(ctx, dtype, N, nnz):
    row, col, val = shuffled_coo(N, nnz, N, dtype)
    C = sp.coo_csr(sp.coo_from_host(N, row, col, val))
    rp, oc, ov = csr_host(C)
    erp, ec, ev = oracle.coo_csr(N, row, col, val)
    assert np.array_equal(rp, erp) and np.array_equal(oc, ec)
    assert np.array_equal(ov.view(np.uint8), ev.view(np.uint8))


def test_coo_csr_large_stable_sort(ctx):
    N, nnz = 1 << 18, 3_000_000
    row, col, val = shuffled_coo(N, nnz, 5, np.float64)
    C = sp.coo_csr(sp.coo_from_host(N, row, col, val))
    rp, oc, ov = csr_host(C)
    order = np.lexsort((np.arange(nnz), col, row))
    assert np.array_equal(oc, col[order]) and np.array_equal(ov, val[order])
    assert np.array_equal(np.diff(rp), np.bincount(row, minlength=N))


def test_coo_csr_empty_and_errors(ctx):
    C = sp.coo_csr(sp.coo_from_host(4, np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0)))
    assert C.row.numpy().tolist() == [0, 0, 0, 0, 0]
    with pytest.raises(SableError) as e:
        sp.coo_csr(sp.coo_from_host(4, np.array([0, 4], np.int32), np.array([0, 0], np.int32), np.ones(2)))
    assert e.value.name == "SABLE_EMALFORMED"


# --------------------------------------------------------------- csr_ell/dia
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_csr_ell_bit_exact(ctx, dtype):
    N = 20000
    rp, c, v = gen.random_csr(N, 4, seed=1, dtype=dtype)
    A = sp.csr_from_host(N, rp, c, v)
    assert sp.num_cols(A) == oracle.num_cols(N, rp)
    E = sp.csr_ell(A)
    nc, idx, data = oracle.csr_ell(N, rp, c, v)
    assert E.ncols == nc
    assert np.array_equal(E.indices.numpy(), idx) and np.array_equal(E.data.numpy(), data)


def test_csr_ell_guard_refuses_like_reference(ctx, capsys):
    rp, c, v = gen.zipf_csr(1 << 14, 1.5, 4096, seed=2)
    A = sp.csr_from_host(len(rp) - 1, rp, c, v)
    assert sp.csr_ell(A) is None
    assert "not efficient to store this matrix in ELL" in capsys.readouterr().err


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("stride_mode", ["N", "N-min"])
def test_csr_dia_bit_exact(ctx, dtype, stride_mode):
    N = 70000
    offs = [-4096, -64, -1, 0, 1, 64, 4096] if stride_mode == "N" else [-9, -3, 5, 11]
    rp, c, v = gen.band_csr(N, offs, seed=3, dtype=dtype)
    A = sp.csr_from_host(N, rp, c, v)
    nd, mn = oracle.num_diags(N, rp, c)
    assert sp.num_diags(A, stride_mode) == [nd, N if stride_mode == "N" else N - mn]
    D = sp.csr_dia(A, stride_mode=stride_mode)
    end, estride, eoff, edata = oracle.csr_dia(N, rp, c, v, stride=D.stride)
    assert D.ndiags == end and D.stride == estride
    assert np.array_equal(D.offset.numpy(), eoff)
    assert np.array_equal(D.data.numpy(), edata)


def test_csr_dia_duplicates_last_wins(ctx):
    N = 5
    rp = np.array([0, 2, 3, 3, 4, 5], np.int32)
    c = np.array([1, 1, 2, 3, 0], np.int32)
    v = np.array([1.0, 2.0, 3.0, 4.0, 5.0])
    A = sp.csr_from_host(N, rp, c, v)
    D = sp.csr_dia(A, guard=False)
    nd, stride, off, data = oracle.csr_dia(N, rp, c, v)
    assert np.array_equal(D.offset.numpy(), off) and np.array_equal(D.data.numpy(), data)


# ---------------------------------------------------------------- transposes
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("dup", [False, True])
def test_transpose_csr_bit_exact(ctx, dtype, dup):
    N = 30000
    rp, c, v = gen.random_csr(N, 6, seed=4, dtype=dtype, duplicates=dup, empty_rows=0.1)
    T = sp.transpose(sp.csr_from_host(N, rp, c, v))
    trp, tc, tv = csr_host(T)
    erp, ec, ev = oracle.transpose_csr(N, rp, c, v)
    assert np.array_equal(trp, erp) and np.array_equal(tc, ec)
    assert np.array_equal(tv.view(np.uint8), ev.view(np.uint8))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_transpose_dia_bit_exact(ctx, dtype):
    N = 50000
    rp, c, v = gen.band_csr(N, [-300, -1, 0, 2, 4096], seed=5, dtype=dtype)
    nd, stride, off, data = oracle.csr_dia(N, rp, c, v)
    D = sp.dia_from_host(N, stride, off, data, len(c))
    T = sp.transpose(D)
    eo, ed = oracle.transpose_dia(N, stride, off, data)
    assert np.array_equal(T.offset.numpy(), eo) and np.array_equal(T.data.numpy(), ed)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_transpose_ell_bit_exact(ctx, dtype):
    N = 40000
    rp, c, v = gen.random_csr(N, 5, seed=6, dtype=dtype)
    v[::5] = 0.0
    v[1::11] = -0.0
    v[2::13] = np.nan
    nc, idx, data = oracle.csr_ell(N, rp, c, v)
    E = sp.ell_from_host(N, nc, idx, data, len(c))
    T = sp.transpose(E)
    nco, ei, ed = oracle.transpose_ell(N, nc, idx, data)
    assert T.ncols == nco
    assert np.array_equal(T.indices.numpy(), ei)
    assert np.array_equal(T.data.numpy().view(np.uint8), ed.view(np.uint8))


def test_transpose_coo_swaps(ctx):
    row, col, val = shuffled_coo(100, 500, 7, np.float64)
    T = sp.transpose(sp.coo_from_host(100, row, col, val))
    assert np.array_equal(T.row.numpy(), col) and np.array_equal(T.col.numpy(), row)
    assert np.array_equal(T.val.numpy(), val)


def test_convert_then_spmv_chain(ctx):
    # COO -> CSR -> ELL / DIA, SpMV in every format agrees (laplacian, exact integers)
    n = 256
    rp, c, v = gen.laplacian_2d(n, np.float32)
    N = n * n
    row = gen.csr_to_coo_rows(rp)
    p = np.random.default_rng(8).permutation(len(row))
    A = sp.coo_csr(sp.coo_from_host(N, row[p], c[p], v[p]))
    assert np.array_equal(A.row.numpy(), rp) and np.array_equal(A.col.numpy(), c)
    x = np.random.default_rng(9).integers(-8, 9, N).astype(np.float32)
    ys = []
    for M in (A, sp.csr_ell(A), sp.csr_dia(A)):
        yv = sp.create_vec(N, 0.0, np.float32)
        sp.spmv(M, sp.vec_from_host(x), yv)
        ys.append(yv.array())
    assert np.array_equal(ys[0], ys[1]) and np.array_equal(ys[0], ys[2])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_coo_csc_bit_exact(ctx, dtype):
    row, col, val = shuffled_coo(3000, 40000, 81, dtype)
    A = sp.coo_from_host(3000, row, col, val)
    C = sp.coo_csc(A)
    cp, r, v = oracle.coo_csc(3000, row, col, val)
    assert np.array_equal(C.col.numpy(), cp)
    assert np.array_equal(C.row.numpy(), r)
    assert np.array_equal(C.val.numpy().view(np.uint8), v.view(np.uint8))
