"""Element-wise ops and eliminate_zeros: the oracle against closed forms.

oracle.ew_values / ew_dia / count_* / copy_nonzeros_* restate the
WebAssembly exports behind seq/src/sparse32.js:200-1207.  They are pinned
here by numpy's IEEE operations (abs, neg, ceil, floor, trunc, rint, sqrt,
multiply, divide), by hand-worked cases for the reference's own rules (sign
of zero and NaN per format, Math.pow's NaN cases, the DIA packing of
copy_nonzeros_dia) and by numpy masks for the compactions.  The
transcendental ops are checked against numpy's double functions rounded to
the value type, which is the reference's widen-call-round sequence.
"""
import numpy as np
import pytest

import oracle

DT = [np.float32, np.float64]


def sample(dtype, n=4000, seed=0):
    rng = np.random.default_rng(seed)
    v = (rng.standard_normal(n) * np.exp(rng.uniform(-5, 5, n))).astype(dtype)
    v[::17] = 0
    v[1::29] = -0.0
    v[2::31] = np.round(v[2::31])
    v[3::37] = np.round(v[3::37]) + 0.5  # ties for nearest
    v[5] = np.nan
    v[6] = np.inf
    v[7] = -np.inf
    return v


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.dtype == b.dtype and np.array_equal(np.isnan(a), np.isnan(b)) and \
        np.array_equal(a[~np.isnan(a)].view(np.uint8), b[~np.isnan(b)].view(np.uint8)) if a.ndim else a == b


@pytest.mark.parametrize("dtype", DT)
def test_oracle_ieee_ops(dtype):
    v = sample(dtype)
    with np.errstate(all="ignore"):
        assert same(oracle.ew_values("abs", 0, v), np.abs(v))
        assert same(oracle.ew_values("neg", 0, v), -v)
        assert same(oracle.ew_values("ceil", 0, v), np.ceil(v))
        assert same(oracle.ew_values("floor", 0, v), np.floor(v))
        assert same(oracle.ew_values("trunc", 0, v), np.trunc(v))
        assert same(oracle.ew_values("nearest", 0, v), np.rint(v))  # half to even
        assert same(oracle.ew_values("sqrt", 0, v), np.sqrt(v))
        assert same(oracle.ew_values("multiply", 2.5, v), v * dtype(2.5))
        p180 = dtype(dtype(np.pi) / dtype(180))
        assert same(oracle.ew_values("deg2rad", np.pi, v), v * p180)
        assert same(oracle.ew_values("rad2deg", np.pi, v), v / p180)


def test_oracle_sign_rules():
    v = np.array([2.0, -3.0, 0.0, -0.0, np.nan], np.float32)
    # COO/CSR exports: v > 0 ? 1 : -1, so 0 and NaN give -1
    assert oracle.ew_values("sign", 0, v).tolist() == [1, -1, -1, -1, -1]
    # DIA/ELL exports: 0 stays 0, NaN gives -1
    assert oracle.ew_values("sign0", 0, v).tolist() == [1, -1, 0, 0, -1]


@pytest.mark.parametrize("dtype", DT)
@pytest.mark.parametrize("op,f", [("expm1", np.expm1), ("log1p", np.log1p), ("sin", np.sin), ("tan", np.tan)])
def test_oracle_transcendental_widen_call_round(dtype, op, f):
    v = sample(dtype)
    with np.errstate(all="ignore"):
        ref = f(v.astype(np.float64)).astype(dtype)
    got = oracle.ew_values(op, 0, v)
    # numpy and the C library may differ in the last bit of the double
    # result; after rounding to f32 they agree, in f64 within 1 ulp
    if dtype == np.float32:
        assert same(got, ref)
    else:
        fin = np.isfinite(ref)
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        assert np.all(np.abs(got[fin].view(np.int64) - ref[fin].view(np.int64)) <= 1)


def test_oracle_pow_follows_math_pow():
    v = np.array([1.0, -1.0, 1.0, 2.0, 0.0, 0.0, -8.0, 4.0], np.float64)
    p = [np.nan, np.inf, np.inf, 0.5, -1.0, 0.0, 1.0 / 3.0, 0.5]
    got = [oracle.ew_values("pow", pi, v[i:i + 1])[0] for i, pi in enumerate(p)]
    # Math.pow(1, NaN) and Math.pow(+-1, +-Infinity) are NaN (C's pow gives 1)
    assert np.isnan(got[0]) and np.isnan(got[1]) and np.isnan(got[2])
    assert got[3] == np.sqrt(2.0) and got[4] == np.inf and got[5] == 1.0
    assert np.isnan(got[6]) and got[7] == 2.0
    v32 = np.array([3.0], np.float32)
    assert oracle.ew_values("pow", 0.5, v32)[0] == np.float32(np.sqrt(3.0))


def test_oracle_dia_in_range_only():
    # N=4, offsets -1, 0, 2 (stride N): padding slots hold 9 and must not change
    N, off = 4, np.array([-1, 0, 2], np.int32)
    data = np.full(12, 9.0, np.float32)
    data[1:4] = [1, 2, 3]      # k=-1: rows 1..3
    data[4:8] = [4, 5, 6, 7]   # k=0
    data[8:10] = [8, 10]       # k=2: rows 0..1
    out = oracle.ew_dia("neg", 0, N, N, off, data, out=data)
    assert out.tolist() == [9, -1, -2, -3, -4, -5, -6, -7, -8, -10, 9, 9]


def test_oracle_copy_nonzeros_dia_packs_like_the_reference():
    # copy_nonzeros_dia keeps diagonals with a nonzero and writes their
    # nonzeros in row order from the slot of the first row, so a zero inside
    # a diagonal moves the later values up one slot
    N, off = 4, np.array([-1, 0, 2], np.int32)
    data = np.zeros(12, np.float32)
    data[1:4] = [1, 0, 3]      # k=-1, rows 1..3, one interior zero
    data[4:8] = [0, 0, 0, 0]   # k=0 all zero: dropped
    data[8:10] = [0, 8]        # k=2, rows 0..1
    assert oracle.count_nonzeros_dia(N, N, off, data) == (2, 3)
    ooff, odata = oracle.copy_nonzeros_dia(N, N, off, data)
    assert ooff.tolist() == [-1, 2]
    assert odata.tolist() == [0, 1, 3, 0, 8, 0, 0, 0]


def test_oracle_compactions_match_masks():
    rng = np.random.default_rng(1)
    N, nnz = 50, 600
    val = rng.standard_normal(nnz).astype(np.float64)
    val[rng.random(nnz) < 0.3] = 0
    val[3] = np.nan
    row = np.sort(rng.integers(0, N, nnz)).astype(np.int32)
    col = rng.integers(0, N, nnz).astype(np.int32)
    keep = (val > 0) | (val < 0)
    assert oracle.count_nonzeros(val) == keep.sum()
    r, c, v = oracle.copy_nonzeros_coo(row, col, val)
    assert np.array_equal(r, row[keep]) and np.array_equal(c, col[keep]) and np.array_equal(v, val[keep])
    rp = np.searchsorted(row, np.arange(N + 1)).astype(np.int32)
    orp, oc, ov = oracle.copy_nonzeros_csr(N, rp, col, val)
    assert np.array_equal(orp, np.searchsorted(row[keep], np.arange(N + 1)))
    assert np.array_equal(oc, col[keep]) and np.array_equal(ov, val[keep])
    # ELL: per-row nonzeros in slot order
    nc = 5
    data = rng.standard_normal(nc * N).astype(np.float32)
    data[rng.random(nc * N) < 0.4] = 0
    idx = rng.integers(0, N, nc * N).astype(np.int32)
    D, I = data.reshape(nc, N), idx.reshape(nc, N)
    m = (D > 0) | (D < 0)
    cc, odata, oidx = oracle.copy_nonzeros_ell(N, nc, data, idx)
    assert cc == m.sum(0).max() and oracle.count_nonzeros_ell(N, nc, data) == (cc, m.sum())
    OD, OI = odata.reshape(cc, N), oidx.reshape(cc, N)
    for i in range(N):
        k = m[:, i].sum()
        assert np.array_equal(OD[:k, i], D[m[:, i], i]) and np.array_equal(OI[:k, i], I[m[:, i], i])
        assert not OD[k:, i].any() and not OI[k:, i].any()
