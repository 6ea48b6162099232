"""Pins the CPU oracle (oracle/sable_oracle.c) against brute force, closed
forms and order-sensitive known answers.  The reference ships no golden
vectors for this path (its only checksums are for SuiteSparse matrices that
are not in the repository), so these checks are what pin each restatement.
CPU only."""
import os

import numpy as np
import pytest

import oracle
from sciwasm.sparse_amd import gen


def dense_from_csr(N, rp, col, val):
    A = np.zeros((N, N), val.dtype)
    for i in range(N):
        for j in range(rp[i], rp[i + 1]):
            A[i, col[j]] = val[j]          # a later duplicate overwrites
    return A


def dense_sum_from_csr(N, rp, col, val):
    A = np.zeros((N, N), np.float64)
    for i in range(N):
        for j in range(rp[i], rp[i + 1]):
            A[i, col[j]] += val[j]
    return A


def random_coo(N, nnz, seed, dtype=np.float64, dup=True):
    rng = np.random.default_rng(seed)
    row = rng.integers(0, N, nnz).astype(np.int32)
    col = rng.integers(0, N, nnz).astype(np.int32)
    if not dup:
        key = np.unique(row.astype(np.int64) * N + col)
        rng.shuffle(key)
        row, col = (key // N).astype(np.int32), (key % N).astype(np.int32)
    val = rng.integers(-9, 10, len(row)).astype(dtype)
    return row, col, val


# ------------------------------------------------------------------ coo_csr
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_coo_csr_is_stable_row_col_sort(dtype, seed):
    N = 40
    row, col, val = random_coo(N, 300, seed, dtype)
    # tie-breaker: make duplicates distinguishable by value
    val = np.arange(len(row)).astype(dtype)
    rp, oc, ov = oracle.coo_csr(N, row, col, val)
    order = np.lexsort((np.arange(len(row)), col, row))  # row, then col, then input position
    assert np.array_equal(oc, col[order])
    assert np.array_equal(ov, val[order])
    assert rp[0] == 0 and rp[-1] == len(row)
    assert np.array_equal(np.diff(rp), np.bincount(row, minlength=N))


def test_coo_csr_empty_rows_and_single():
    rp, oc, ov = oracle.coo_csr(5, np.array([3], np.int32), np.array([1], np.int32), np.array([2.5]))
    assert rp.tolist() == [0, 0, 0, 0, 1, 1]
    assert oc.tolist() == [1] and ov.tolist() == [2.5]


# --------------------------------------------------------------- csr_ell/dia
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_csr_ell_dense_and_layout(dtype):
    N = 30
    rp, c, v = gen.random_csr(N, 4, seed=3, dtype=dtype)
    nc, idx, data = oracle.csr_ell(N, rp, c, v)
    assert nc == np.diff(rp).max()
    D = np.zeros((N, N), dtype)
    for k in range(nc):
        for i in range(N):
            D[i, idx[k * N + i]] += data[k * N + i]
    assert np.array_equal(D, dense_from_csr(N, rp, c, v))
    # slot k of row i holds the k-th stored entry; padding is (0, 0)
    for i in range(N):
        L = rp[i + 1] - rp[i]
        assert np.array_equal(idx[np.arange(L) * N + i], c[rp[i]:rp[i + 1]])
        assert np.all(idx[np.arange(L, nc) * N + i] == 0)
        assert np.all(data[np.arange(L, nc) * N + i] == 0)


def test_csr_dia_laplacian_closed_form():
    n = 16
    rp, c, v = gen.laplacian_2d(n)
    N = n * n
    nd, mn = oracle.num_diags(N, rp, c)
    assert (nd, mn) == (5, 0)
    nd, stride, off, data = oracle.csr_dia(N, rp, c, v)
    assert off.tolist() == [-n, -1, 0, 1, n]
    assert stride == N
    D = data.reshape(nd, N)
    assert np.all(D[2] == 4.0)
    # row i of the offset -1 diagonal holds a(i, i-1): -1 unless i is at a grid row start
    assert np.array_equal(D[1], np.where(np.arange(N) % n == 0, 0.0, -1.0))
    assert np.array_equal(D[0], np.where(np.arange(N) < n, 0.0, -1.0))


@pytest.mark.parametrize("stride_mode", ["N", "N-min"])
def test_csr_dia_dense_roundtrip(stride_mode):
    N = 50
    rp, c, v = gen.band_csr(N, [-7, -3, 2, 5], seed=4)
    nd, mn = oracle.num_diags(N, rp, c)
    assert (nd, mn) == (4, 2)
    stride = N if stride_mode == "N" else N - mn
    nd, stride, off, data = oracle.csr_dia(N, rp, c, v, stride=stride)
    D = np.zeros((N, N))
    for k, o in enumerate(off):
        for i in range(max(0, -o), min(N, N - o)):
            D[i, i + o] = data[k * stride + i - (N - stride if o < 0 else 0)]
    assert np.array_equal(D, dense_from_csr(N, rp, c, v))


# ---------------------------------------------------------------- transposes
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_transpose_csr(dtype):
    N = 37
    rp, c, v = gen.random_csr(N, 5, seed=5, dtype=dtype)
    trp, tc, tv = oracle.transpose_csr(N, rp, c, v)
    assert np.array_equal(dense_from_csr(N, trp, tc, tv), dense_from_csr(N, rp, c, v).T)
    for i in range(N):
        assert np.all(np.diff(tc[trp[i]:trp[i + 1]]) > 0)
    rrp, rc, rv = oracle.transpose_csr(N, trp, tc, tv)
    assert np.array_equal(rrp, rp) and np.array_equal(rc, c) and np.array_equal(rv, v)


def test_transpose_csr_duplicates_keep_row_order():
    N = 4
    rp = np.array([0, 2, 3, 3, 4], np.int32)
    c = np.array([1, 1, 1, 1], np.int32)
    v = np.array([10.0, 11.0, 20.0, 40.0])
    trp, tc, tv = oracle.transpose_csr(N, rp, c, v)
    assert trp.tolist() == [0, 0, 4, 4, 4]
    assert tc.tolist() == [0, 0, 1, 3] and tv.tolist() == [10.0, 11.0, 20.0, 40.0]


def test_transpose_dia():
    N = 40
    rp, c, v = gen.band_csr(N, [-5, 0, 1, 9], seed=6)
    nd, stride, off, data = oracle.csr_dia(N, rp, c, v)
    toff, tdata = oracle.transpose_dia(N, stride, off, data)
    assert toff.tolist() == [-o for o in off]          # descending
    D = np.zeros((N, N))
    for k, o in enumerate(toff):
        for i in range(max(0, -o), min(N, N - o)):
            D[i, i + o] = tdata[k * stride + i]
    assert np.array_equal(D, dense_from_csr(N, rp, c, v).T)


def test_transpose_ell_drops_zeros():
    N = 25
    rp, c, v = gen.random_csr(N, 3, seed=7)
    v[::4] = 0.0
    nc, idx, data = oracle.csr_ell(N, rp, c, v)
    nco, tidx, tdata = oracle.transpose_ell(N, nc, idx, data)
    A = dense_from_csr(N, rp, c, v)
    assert nco == (A != 0).sum(0).max()
    D = np.zeros((N, N))
    for k in range(nco):
        for i in range(N):
            if tdata[k * N + i] != 0:
                D[i, tidx[k * N + i]] = tdata[k * N + i]
    assert np.array_equal(D, A.T)
    for i in range(N):   # new row entries ascend by column
        cols = [tidx[k * N + i] for k in range(nco) if tdata[k * N + i] != 0]
        assert cols == sorted(cols)


# ---------------------------------------------------------------- level sets
def test_level_sets_laplacian_3d_closed_form():
    n = 6
    rp, c, v = gen.laplacian_3d_lower(n)
    N = n ** 3
    nl, lp, perm, order, lvl = oracle.level_sets(N, rp, c)
    i = np.arange(N)
    gx, gy, gz = i % n, (i // n) % n, i // (n * n)
    assert nl == 3 * (n - 1) + 1
    assert np.array_equal(lvl, gx + gy + gz)
    sizes = np.bincount(gx + gy + gz)
    assert np.array_equal(np.diff(lp), sizes)
    # trinomial count of (i,j,k) in [0,n)^3 with i+j+k = l
    for l in range(nl):
        cnt = sum(1 for a in range(n) for b in range(n) if 0 <= l - a - b < n)
        assert sizes[l] == cnt
    assert np.array_equal(perm, np.lexsort((i, lvl)))
    assert np.array_equal(order[perm], np.arange(N))


def test_level_sets_random_lower_invariants():
    N = 300
    rp, c, v = gen.random_lower(N, 8, seed=8)
    nl, lp, perm, order, lvl = oracle.level_sets(N, rp, c)
    for i in range(N):
        deps = c[rp[i]:rp[i + 1] - 1]
        if len(deps) == 0:
            assert lvl[i] == 0
        else:
            assert lvl[i] == 1 + lvl[deps].max()
    assert sorted(perm.tolist()) == list(range(N))
    nrp, nc_, nv = oracle.level_reorder(N, rp, c, v, perm, order)
    # reordered matrix is the permuted one: P A P^T
    A = dense_from_csr(N, rp, c, v)
    B = dense_from_csr(N, nrp, nc_, nv)
    assert np.array_equal(B, A[np.ix_(perm, perm)])


def test_level_sets_empty_row_is_level_one():
    rp = np.array([0, 1, 1, 3], np.int32)
    c = np.array([0, 0, 2], np.int32)
    nl, lp, perm, order, lvl = oracle.level_sets(3, rp, c)
    assert lvl.tolist() == [0, 1, 1] and nl == 2 and lp.tolist() == [0, 1, 3]


# ---------------------------------------------------------------------- spmv
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_spmv_exact_integer_all_formats(dtype):
    N = 60
    rp, c, v = gen.random_csr(N, 6, seed=9, dtype=dtype)
    v = np.round(v * 4).astype(dtype)
    x = np.random.default_rng(1).integers(-5, 6, N).astype(dtype)
    y0 = np.random.default_rng(2).integers(-5, 6, N).astype(dtype)
    exact = dense_sum_from_csr(N, rp, c, v) @ x.astype(np.float64) + y0
    assert np.array_equal(oracle.spmv_csr(N, rp, c, v, x, y0), exact.astype(dtype))
    row = gen.csr_to_coo_rows(rp)
    assert np.array_equal(oracle.spmv_coo(row, c, v, x, y0), exact.astype(dtype))
    nc, idx, data = oracle.csr_ell(N, rp, c, v)
    assert np.array_equal(oracle.spmv_ell(N, nc, idx, data, x, y0), exact.astype(dtype))
    nd, stride, off, dd = oracle.csr_dia(N, rp, c, v)
    assert np.array_equal(oracle.spmv_dia(N, stride, off, dd, x, y0), exact.astype(dtype))
    # inner_max repeats y += A.x
    y3 = oracle.spmv_csr(N, rp, c, v, x, np.zeros(N, dtype), inner_max=3)
    assert np.array_equal(y3, (3 * (exact - y0)).astype(dtype))


@pytest.mark.parametrize("dtype,big", [(np.float64, 1e16), (np.float32, 1e8)])
def test_spmv_sequential_order(dtype, big):
    # y = ((0 + big) + 1) + (-big): sequential order loses the 1; exact is 1.
    rp = np.array([0, 3], np.int32)
    c = np.array([0, 0, 0], np.int32)
    v = np.array([big, 1.0, -big], dtype)
    x = np.ones(1, dtype)
    assert oracle.spmv_csr(1, rp, c, v, x, np.zeros(1, dtype))[0] == 0.0
    assert oracle.spmv_coo(np.zeros(3, np.int32), c, v, x, np.zeros(1, dtype))[0] == 0.0
    v2 = np.array([1.0, big, -big], dtype)       # (1 + big) - big = 0 too, but
    v3 = np.array([big, -big, 1.0], dtype)       # (big - big) + 1 = 1
    assert oracle.spmv_csr(1, rp, c, v2, x, np.zeros(1, dtype))[0] == 0.0
    assert oracle.spmv_csr(1, rp, c, v3, x, np.zeros(1, dtype))[0] == 1.0


def test_spmv_laplacian_closed_forms():
    n = 32
    rp, c, v = gen.laplacian_2d(n)
    N = n * n
    y = oracle.spmv_csr(N, rp, c, v, np.ones(N), np.zeros(N))
    i = np.arange(N)
    nb = (i % n > 0).astype(int) + (i % n < n - 1) + (i // n > 0) + (i // n < n - 1)
    assert np.array_equal(y, 4.0 - nb)
    y = oracle.spmv_csr(N, rp, c, v, i.astype(np.float64), np.zeros(N))
    interior = (i % n > 0) & (i % n < n - 1) & (i // n > 0) & (i // n < n - 1)
    assert np.all(y[interior] == 0)


def test_spmv_fp64_oracle():
    N = 50
    rp, c, v = gen.random_csr(N, 5, seed=10)
    x = np.random.default_rng(3).uniform(-1, 1, N)
    y, ab = oracle.spmv_csr_fp64(N, rp, c, v, x)
    assert np.allclose(y, dense_sum_from_csr(N, rp, c, v) @ x, rtol=1e-14, atol=1e-14)
    assert np.all(ab >= np.abs(y) - 1e-15)


# ---------------------------------------------------------------------- spts
def test_spts_manufactured_exact():
    N = 200
    rng = np.random.default_rng(11)
    rp, c, v = gen.random_lower(N, 6, seed=11)
    v = np.where(c == np.repeat(np.arange(N), np.diff(rp)), 1.0, rng.integers(-2, 3, len(c))).astype(np.float64)
    xs = rng.integers(-3, 4, N).astype(np.float64)
    b = dense_sum_from_csr(N, rp, c, v) @ xs
    assert np.array_equal(oracle.spts_syncfree(N, rp, c, v, b), xs)
    assert np.array_equal(oracle.spts_fp64(N, rp, c, v, b), xs)
    nl, lp, perm, order, lvl = oracle.level_sets(N, rp, c)
    nrp, ncc, nv = oracle.level_reorder(N, rp, c, v, perm, order)
    yl = oracle.spts_level(N, nrp, ncc, nv, b[perm])
    assert np.array_equal(yl, xs[perm])
    assert oracle.rel_residual(N, rp, c, v, xs, b) == 0.0


def test_spts_summation_orders_differ_as_specified():
    # row 2: (b - (0 + 1e16*y0 + 1*y1)) / 1  vs  ((b - 1e16*y0) - 1*y1) / 1
    rp = np.array([0, 1, 2, 5], np.int32)
    c = np.array([0, 1, 0, 1, 2], np.int32)
    v = np.array([1.0, 1.0, 1e16, 1.0, 1.0])
    b = np.array([1.0, 1.0, 1e16])
    ys = oracle.spts_syncfree(3, rp, c, v, b)
    assert ys.tolist() == [1.0, 1.0, -1.0]           # (1e16 - 1e16) - 1
    yl = oracle.spts_level(3, rp, c, v, b)
    assert yl.tolist() == [1.0, 1.0, 0.0]            # 1e16 - (1e16 + 1) rounds to 0


def test_fletcher_known_answers():
    assert oracle.fletcher_sum(np.array([1, 2, 3], np.float32)) == 2566.0
    y = np.array([300.0, -1.5], np.float64)
    # s1 = 45, s2 = 45; s1 = 43.5, s2 = 88.5 -> 88.5*256 + 43.5 = 22699.5 -> 22699
    assert oracle.fletcher_sum(y) == 22699.0


def test_static_nnz_partition_known_answer():
    # row lengths 5,1,1,1,1,1,5,1 over 3 workers: ideal 16//3 = 5 -> rows [0,1);
    # then 11//2 = 5 -> rows [1,6); the last worker takes the rest
    lens = np.array([5, 1, 1, 1, 1, 1, 5, 1])
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    rs, re = oracle.static_nnz(8, rp, 3)
    assert list(rs) == [0, 1, 6] and list(re) == [1, 6, 8]
    # ideal 4//4 = 1 -> row 0; then ideal 1//3 = 0 -> empty ranges; the last
    # worker (ideal 1//1) takes row 1
    rs, re = oracle.static_nnz(2, np.array([0, 3, 4], np.int32), 4)
    assert list(rs) == [0, 1, 1, 1] and list(re) == [1, 1, 1, 2]


def test_threaded_cpu_baseline_matches_serial():
    rp, col, val = gen.random_csr(5000, 9, seed=3, empty_rows=0.1)
    x = np.random.default_rng(4).uniform(-1, 1, 5000)
    _, y1 = oracle.time_spmv_csr_f64(5000, rp, col, val, x, warm=0, reps=1)
    _, y4 = oracle.time_spmv_csr_f64_mt(5000, rp, col, val, x, 4, warm=0, reps=2)
    assert np.array_equal(y1, y4)


def lower_csc_from_csr(N, rp, col, val):
    """CSC of a lower CSR (diagonal last in each row) with the diagonal first
    in each column: entries sorted by (column, row)."""
    rows = np.repeat(np.arange(N), np.diff(rp))
    order = np.lexsort((rows, col))
    cp = np.zeros(N + 1, np.int32)
    np.add.at(cp, col + 1, 1)
    return np.cumsum(cp).astype(np.int32), rows[order].astype(np.int32), val[order]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_spts_csc_equals_row_order_and_manufactured(dtype):
    """The column sweep performs, per row, the subtractions in ascending
    column order and then the division: the same operations, in the same
    order, as the sync-free row order on the CSR of the same matrix, so the
    two must agree bit for bit; on an integer manufactured solution both are
    exact."""
    rp, col, val = gen.random_lower(400, kmax=12, seed=5, dtype=dtype)
    N = len(rp) - 1
    cp, row, cval = lower_csc_from_csr(N, rp, col, val)
    b = np.random.default_rng(6).uniform(-1, 1, N).astype(dtype)
    assert np.array_equal(oracle.spts_csc(N, cp, row, cval, b), oracle.spts_syncfree(N, rp, col, val, b))
    # integer L with unit diagonal and integer x: b = L x is exact, so is x
    rng = np.random.default_rng(7)
    ival = rng.integers(-3, 4, len(col)).astype(dtype)
    ival[rp[1:] - 1] = 1
    x = rng.integers(-5, 6, N).astype(dtype)
    bx = np.zeros(N, np.float64)
    rows = np.repeat(np.arange(N), np.diff(rp))
    np.add.at(bx, rows, ival.astype(np.float64) * x[col])
    cp2, row2, cval2 = lower_csc_from_csr(N, rp, col, ival)
    if np.max(np.abs(bx)) < 2 ** 20:
        assert np.array_equal(oracle.spts_csc(N, cp2, row2, cval2, bx.astype(dtype)), x)


def test_coo_csc_is_stable_col_row_sort():
    rng = np.random.default_rng(8)
    N, nnz = 50, 400
    row = rng.integers(0, N, nnz).astype(np.int32)
    col = rng.integers(0, N, nnz).astype(np.int32)
    val = rng.standard_normal(nnz)
    cp, r, v = oracle.coo_csc(N, row, col, val)
    order = np.lexsort((np.arange(nnz), row, col))
    assert np.array_equal(r, row[order]) and np.array_equal(v, val[order])
    assert np.array_equal(cp, np.concatenate([[0], np.cumsum(np.bincount(col, minlength=N))]))


# ------------------------------------------------------------------ row reordering
def test_sort_rows_by_nnz_hand():
    rp = np.array([0, 3, 3, 4, 6, 9], np.int32)
    col = np.arange(9, dtype=np.int32)
    val = np.arange(9, dtype=np.float64)
    orp, oc, ov, perm, nnz_row, bins = oracle.sort_rows_by_nnz(5, rp, col, val)
    # lengths 3,0,1,2,3 -> rows 1,2,3,0,4 (ties keep their order)
    assert perm.tolist() == [1, 2, 3, 0, 4]
    assert orp.tolist() == [0, 0, 1, 3, 6, 9] and oc.tolist() == [3, 4, 5, 0, 1, 2, 6, 7, 8]
    assert nnz_row.tolist() == [3, 0, 1, 2, 3] and bins.tolist() == [1, 2, 3, 5]


def test_sort_rows_by_nnz_is_a_stable_sort_and_spmv_invariant():
    from sciwasm.sparse_amd import gen
    rp, col, val = gen.zipf_csr(3000, 1.4, 64, seed=3)
    N = len(rp) - 1
    orp, oc, ov, perm, nnz_row, bins = oracle.sort_rows_by_nnz(N, rp, col, val)
    lens = np.diff(rp)
    assert np.array_equal(perm, np.argsort(lens, kind="stable"))
    assert np.array_equal(np.diff(orp), lens[perm])
    assert bins.tolist() == [int((lens <= k).sum()) for k in range(4)]
    x = np.random.default_rng(0).standard_normal(N)
    z = np.zeros(N)
    y = oracle.spmv_csr(N, rp, col, val, x, z, 1)
    ys = oracle.spmv_csr(N, orp, oc, ov, x, z, 1)
    # each row keeps its entries in order, so each product is bit-identical
    assert np.array_equal(oracle.sort_y_rows_by_nnz(perm, ys), y)


def test_reorder_nn_hand():
    """64 rows, rows 0-3 hold one entry each in cache lines 0, 2, 0, 2
    (columns 0, 40, 5, 45); the rest are empty.  After the sort by nnz they
    are rows 60-63.  Worked by hand through reorder_subset: 60 starts; its
    list gets 61 (similarity 0), then 62 (1.0) overwrites 61 in place (the
    insertion replaces the last entry instead of shifting it), then 63 (0);
    62 follows; from 62 the tie between 63 (slot 0) and 61 (slot 1) goes to
    the later slot, 61; then 63."""
    N = 64
    rp = np.zeros(N + 1, np.int32)
    rp[1:5] = [1, 2, 3, 4]
    rp[5:] = 4
    col = np.array([0, 40, 5, 45], np.int32)
    val = np.ones(4, np.float64)
    orp, oc, _, sperm, _, _ = oracle.sort_rows_by_nnz(N, rp, col, val)
    assert sperm[60:].tolist() == [0, 1, 2, 3]
    perm = oracle.reorder_nn(N, orp, oc, 16)
    assert perm[:60].tolist() == list(range(60))
    assert perm[60:].tolist() == [60, 62, 61, 63]
    assert sperm[perm][60:].tolist() == [0, 2, 1, 3]


def test_reorder_nn_is_a_permutation_that_groups_shared_lines():
    from sciwasm.sparse_amd import gen
    rp, col, val = gen.laplacian_2d(48)  # 2304 rows, 5-point stencil
    N = len(rp) - 1
    # scatter the rows, so that neighbouring grid rows are far apart
    p = np.random.default_rng(0).permutation(N)
    lens = np.diff(rp)[p]
    srp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    scol = np.concatenate([col[rp[r]:rp[r + 1]] for r in p]).astype(np.int32)
    orp, oc, ov, sperm, _, _ = oracle.sort_rows_by_nnz(N, srp, scol, np.ones(len(scol)))
    perm = oracle.reorder_nn(N, orp, oc, 16)
    assert np.array_equal(np.sort(perm), np.arange(N))

    def shared(rpx, colx, order):
        # cache lines shared by consecutive rows of the ordering
        s = 0
        for a, b in zip(order[:-1], order[1:]):
            la = set((colx[rpx[a]:rpx[a + 1]] // 16).tolist())
            lb = set((colx[rpx[b]:rpx[b + 1]] // 16).tolist())
            s += len(la & lb)
        return s
    assert shared(orp, oc, perm) > 2 * shared(orp, oc, np.arange(N))


# ------------------------------------------------------------------ row-major layouts
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_rowmajor_conversions_are_the_transposed_layouts(dtype):
    rng = np.random.default_rng(11)
    rp, col, _ = gen.laplacian_2d(9)
    N = len(rp) - 1
    val = rng.uniform(-1, 1, len(col)).astype(dtype)
    nc, ci, cd = oracle.csr_ell(N, rp, col, val)
    nr, ri, rd = oracle.csr_ell_row(N, rp, col, val)
    assert nr == nc == 5
    assert np.array_equal(ri.reshape(N, nc), ci.reshape(nc, N).T)
    assert np.array_equal(rd.reshape(N, nc), cd.reshape(nc, N).T)
    nd, _, off, dd = oracle.csr_dia(N, rp, col, val)
    nd2, off2, rdd = oracle.csr_dia_row(N, rp, col, val)
    assert nd2 == nd == 5 and np.array_equal(off2, off) and list(off) == [-9, -1, 0, 1, 9]
    assert np.array_equal(rdd.reshape(N, nd), dd.reshape(nd, N).T)
    # dense round trip of the row-major DIA
    A = dense_from_csr(N, rp, col, val)
    B = np.zeros_like(A)
    for i in range(N):
        for k in range(nd):
            c = i + off2[k]
            if 0 <= c < N:
                B[i, c] = rdd[i * nd + k]
    assert np.array_equal(A, B)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_rowmajor_spmv_equals_column_major_bits(dtype):
    rng = np.random.default_rng(12)
    rp, col, _ = gen.laplacian_2d(12)
    N = len(rp) - 1
    val = rng.uniform(-1, 1, len(col)).astype(dtype)
    x = rng.uniform(-1, 1, N).astype(dtype)
    y0 = rng.uniform(-1, 1, N).astype(dtype)
    nd, _, off, dd = oracle.csr_dia(N, rp, col, val)
    _, _, rdd = oracle.csr_dia_row(N, rp, col, val)
    yc = oracle.spmv_dia(N, N, off, dd, x, y0, 3)
    yr = oracle.spmv_dia_row(N, off, rdd, x, y0, inner_max=3)
    assert np.array_equal(yc.view(np.uint8), yr.view(np.uint8))
    nc, ci, cd = oracle.csr_ell(N, rp, col, val)
    _, ri, rd = oracle.csr_ell_row(N, rp, col, val)
    ye = oracle.spmv_ell(N, nc, ci, cd, x, y0, 2)
    yer = oracle.spmv_ell_row(N, nc, ri, rd, x, y0, inner_max=2)
    assert np.array_equal(ye.view(np.uint8), yer.view(np.uint8))
    # a row range touches only its rows
    part = oracle.spmv_dia_row(N, off, rdd, x, y0, r0=20, r1=50)
    full = oracle.spmv_dia_row(N, off, rdd, x, y0)
    assert np.array_equal(part[20:50], full[20:50])
    assert np.array_equal(part[:20], y0[:20]) and np.array_equal(part[50:], y0[50:])


def test_rowmajor_ell_stops_at_negative_index_and_gs_order():
    f = np.float32
    # plain order leaves the row at the first negative index
    idx = np.array([0, 1, -1, 2], np.int32)
    data = np.array([1, 2, 4, 8], f)
    x = np.array([1, 1, 1], f)
    y = oracle.spmv_ell_row(1, 4, idx, data, x, np.zeros(1, f))
    assert y[0] == 3
    # gs order, width 8: lanes take slots k and k + 4; with s0 = 2^24,
    # s1 = 1, s4 = -2^24 the lane sums are (0, 1, 0, 0), total 1, where the
    # sequential order rounds 2^24 + 1 to 2^24 and ends at 0
    idx = np.zeros(8, np.int32)
    data = np.array([2.0 ** 24, 1, 0, 0, -2.0 ** 24, 0, 0, 0], f)
    x1 = np.ones(1, f)
    assert oracle.spmv_ell_row(1, 8, idx, data, x1, np.zeros(1, f), gs=True)[0] == 1
    assert oracle.spmv_ell_row(1, 8, idx, data, x1, np.zeros(1, f), gs=False)[0] == 0
    # width 5: slot 0 is the head, added to y before the lanes
    data5 = np.array([1, 2.0 ** 24, 1, 0, 0], f)
    y5 = oracle.spmv_ell_row(1, 5, np.zeros(5, np.int32), data5, x1, np.array([2.0 ** 24], f), gs=True)
    # head: 2^24 + 1 -> 2^24; lanes (2^24, 1, 0, 0): ((2^24 + 2^24) + 1) -> 2^25
    assert y5[0] == 2.0 ** 25
    # width < 4 takes the plain order
    d3 = np.array([1, 2, 3], f)
    assert oracle.spmv_ell_row(1, 3, np.zeros(3, np.int32), d3, x1, np.zeros(1, f), gs=True)[0] == 6
    # exact integer products agree with the dense product in both orders
    rng = np.random.default_rng(5)
    N, nc = 40, 6
    ci = rng.integers(0, N, N * nc).astype(np.int32)
    cd = rng.integers(-5, 6, N * nc).astype(np.float64)
    xv = rng.integers(-3, 4, N).astype(np.float64)
    ref = (cd.reshape(N, nc) * xv[ci.reshape(N, nc)]).sum(1)
    for gs in (False, True):
        assert np.array_equal(oracle.spmv_ell_row(N, nc, ci, cd, xv, np.zeros(N), gs=gs), ref)


def test_threaded_baseline_under_tsan():
    """The threaded CPU baseline (persistent workers over static_nnz row
    ranges, barriers around each product) built with ThreadSanitizer: no
    race reported, and every worker count gives the serial restatement's
    bits."""
    import shutil
    import subprocess
    if shutil.which("make") is None or shutil.which("gcc") is None:
        pytest.skip("make or gcc missing")
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    r = subprocess.run(["make", "-s", "-C", here, "tsan"], capture_output=True, text=True, timeout=300)
    if "unsupported" in r.stderr.lower() or "cannot find -ltsan" in r.stderr:
        pytest.skip("no ThreadSanitizer runtime: " + r.stderr[-200:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ThreadSanitizer" not in r.stdout + r.stderr
    assert "equal to the serial restatement" in r.stdout
