"""Row reordering on the GPU against the oracle (par/src/sparse32.js:
2148-2235): permutation, row_ptr, col, val, nnz per row and the short-row
counts bit-exact; a product with the reordered matrix, put back in row
order, within the CSR summation bound of the product with the original."""
import numpy as np
import pytest

import oracle
import sciwasm.sparse_amd as sp
from sciwasm.sparse_amd import gen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("kind", ["zipf", "empty_rows", "single", "duplicates"])
def test_sort_rows_by_nnz_matches_oracle(ctx, dtype, kind):
    if kind == "zipf":
        rp, col, val = gen.zipf_csr(200000, 1.5, 512, seed=1, dtype=dtype)
    elif kind == "empty_rows":
        rp, col, val = gen.random_csr(50000, 3, seed=2, dtype=dtype, empty_rows=0.3)
    elif kind == "single":
        rp, col, val = np.array([0, 1], np.int32), np.array([0], np.int32), np.array([2], dtype)
    else:
        # a row longer than N (duplicate entries)
        rp, col, val = np.array([0, 0, 3], np.int32), np.array([1, 1, 0], np.int32), np.array([1, 2, 3], dtype)
    N = len(rp) - 1
    A = sp.csr_from_host(N, rp, col, val, ctx)
    B = sp.sort_rows_by_nnz(A)
    orp, oc, ov, perm, nnz_row, bins = oracle.sort_rows_by_nnz(N, rp, col, val)
    assert np.array_equal(B.permutation.numpy(), perm)
    assert np.array_equal(B.row.numpy(), orp) and np.array_equal(B.col.numpy(), oc)
    assert np.array_equal(B.val.numpy().view(np.uint8), ov.view(np.uint8))
    assert np.array_equal(B.nnz_row.numpy(), nnz_row)
    assert [B.num_zero_rows, B.num_one_rows, B.num_two_rows, B.num_three_rows] == bins.tolist()
    # product on the reordered matrix, back in the original row order
    x = np.random.default_rng(3).standard_normal(max(N, int(col.max()) + 1)).astype(dtype)
    y = sp.create_vec(N, 0.0, dtype, ctx)
    sp.spmv(B, sp.vec_from_host(x, ctx=ctx), y)
    sp.sort_y_rows_by_nnz(y, B)
    ref = oracle.spmv_csr(N, rp, col, val, x, np.zeros(N, dtype), 1)
    u = np.finfo(dtype).eps / 2
    k = np.diff(rp).astype(np.float64) + 1
    rows = np.repeat(np.arange(N), np.diff(rp))
    bound = 2 * (k * u / (1 - k * u)) * np.bincount(rows, np.abs(val.astype(np.float64) * x[col]), N) + \
        np.finfo(dtype).tiny
    assert np.all(np.abs(y.vec.numpy().astype(np.float64) - ref) <= bound)


def test_permute_rows_round_trip(ctx):
    N = 100000
    perm = np.random.default_rng(4).permutation(N).astype(np.int32)
    x = np.random.default_rng(5).standard_normal(N)
    dp = sp.DeviceArray.from_host(perm, ctx=ctx)
    dx = sp.DeviceArray.from_host(x, ctx=ctx)
    g = sp.permute_rows(dx, dp)
    assert np.array_equal(g.numpy(), x[perm])
    back = sp.permute_rows(g, dp, scatter=True)
    assert np.array_equal(back.numpy(), x)


def scattered_laplacian(n, seed):
    rp, col, val = gen.laplacian_2d(n)
    N = len(rp) - 1
    p = np.random.default_rng(seed).permutation(N)
    lens = np.diff(rp)[p]
    srp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    idx = np.concatenate([np.arange(rp[r], rp[r + 1]) for r in p])
    return N, srp, col[idx].astype(np.int32), val[idx]


@pytest.mark.parametrize("w", [16, 4, 1])
@pytest.mark.parametrize("kind", ["laplacian", "random", "zipf"])
def test_reorder_nn_matches_oracle(ctx, kind, w):
    if kind == "laplacian":
        N, rp, col, val = scattered_laplacian(150, 1)  # 22500 rows, buckets of 1024 and remainders
    elif kind == "random":
        rp, col, val = gen.random_csr(30000, 3, seed=2, dtype=np.float64, empty_rows=0.1)
        N = len(rp) - 1
    else:
        rp, col, val = gen.zipf_csr(40000, 1.6, 256, seed=3)
        N = len(rp) - 1
    print("reorder_nn", kind, w, N, flush=True)
    A = sp.csr_from_host(N, rp, col, val, ctx)
    S = sp.sort_rows_by_nnz(A)
    R = sp.reorder_NN(S, w)
    orp, oc, ov, sperm, _, _ = oracle.sort_rows_by_nnz(N, rp, col, val)
    pnn = oracle.reorder_nn(N, orp, oc, w)
    final = sperm[pnn]
    assert np.array_equal(R.permutation.numpy(), final)
    assert np.array_equal(R.row.numpy(), np.concatenate([[0], np.cumsum(np.diff(rp)[final])]))
    assert np.array_equal(R.col.numpy(), np.concatenate([col[rp[r]:rp[r + 1]] for r in final]))
    assert R.num_one_rows == S.num_one_rows
