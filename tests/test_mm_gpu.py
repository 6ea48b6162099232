"""MatrixMarket ingest on the GPU against the oracle (seq/src/sparse32.js:
1889-2172 restated in oracle/sable_oracle.c).

Bit-exact: indices and the value bits of both COO variants (full and
lower), and the mm_info sizes.  Files use distinct (row, col) keys, for
which the reference's quicksort order is the (row, col) order; formats the
reference misreads (tabs, runs of spaces, CRLF, blank lines, a last line
without '\\n') are compared with the oracle on the same file in the
reference's own layout.
"""
import numpy as np
import pytest

import oracle
import sciwasm.sparse_amd as sp
from sciwasm.sparse_amd import gen
from sciwasm.sparse_amd._lib import SableError

from test_mm import decimal_cases, random_mm

pytestmark = pytest.mark.gpu


def coo_host(A):
    return A.row.numpy(), A.col.numpy(), A.val.numpy()


def check_against_oracle(text, dtype, ctx, lower=False):
    info = sp.read_matrix_MM_files(text, ctx)
    A = (sp.create_LCOO_from_MM if lower else sp.create_COO_from_MM)(info, dtype)
    oi, r, c, v = oracle.mm_coo(text, dtype, lower=lower)
    assert (info.nrows, info.ncols, info.nentries) == (oi["nrows"], oi["ncols"], oi["nentries"])
    assert A.nnz == oi["nnz"] == (info.lower_nnz if lower else info.nnz)
    gr, gc, gv = coo_host(A)
    assert np.array_equal(gr, r)
    assert np.array_equal(gc, c)
    assert np.array_equal(gv.view(np.uint8), v.view(np.uint8))
    return info, A


@pytest.mark.parametrize("field,symmetry", [("real", "general"), ("real", "symmetric"), ("integer", "general"),
                                            ("pattern", "general"), ("pattern", "symmetric"),
                                            ("complex", "general")])
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("lower", [False, True], ids=["full", "lower"])
def test_mm_matches_oracle(ctx, field, symmetry, dtype, lower):
    print("mm", field, symmetry, np.dtype(dtype).name, lower, flush=True)
    if field == "complex":
        # "re im" per entry; the reference keeps the real part
        text = random_mm(3, "real", "general", n=80, nent=900)
        lines = text.split(b"\n")
        head, body = lines[:3], [ln + b" 0.5" for ln in lines[3:] if ln]
        text = b"\n".join([head[0].replace(b"real", b"complex")] + head[1:] + body) + b"\n"
    else:
        text = random_mm(2, field, symmetry, n=80, nent=900)
    check_against_oracle(text, dtype, ctx, lower)


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_mm_large_matches_oracle(ctx, dtype):
    """1M distinct entries, 17-digit values, symmetric: several LDS-staged
    workgroups, many line-head tiles, the full radix sort."""
    rng = np.random.default_rng(5)
    n, nent = 200000, 1000000
    r = rng.integers(0, n, int(nent * 1.1))
    c = rng.integers(0, n, len(r))
    r, c = np.maximum(r, c), np.minimum(r, c)  # lower triangle, as symmetric files store it
    _, first = np.unique(r * n + c, return_index=True)
    keep = rng.permutation(first)[:nent]  # distinct keys in random file order
    r, c = r[keep] + 1, c[keep] + 1
    pool = gen.value_pool(4096, seed=1, style="e16")
    text = gen.mm_text(n, n, r, c, pool, rng.integers(0, len(pool), len(r)), symmetry="symmetric")
    print("large mm", len(text), "bytes", flush=True)
    info, A = check_against_oracle(text, dtype, ctx)
    assert info.nnz > len(r)
    check_against_oracle(text, dtype, ctx, lower=True)


def test_mm_values_round_like_float(ctx):
    """Every value of a one-entry-per-row file converts to float(token)'s
    bits: ties, subnormals, overflow, 1-500 digit tokens."""
    toks = [t for t in decimal_cases(21, 20000, 2000)]
    n = len(toks)
    text = gen.mm_text(n, 1, np.arange(1, n + 1), np.ones(n, np.int64), toks, np.arange(n))
    print("value stress", n, flush=True)
    info = sp.read_matrix_MM_files(text, ctx)
    ref = np.array([float(t) for t in toks])
    keep = ref != 0
    A = sp.create_COO_from_MM(info, np.float64)
    r, _, v = coo_host(A)
    assert np.array_equal(r, np.nonzero(keep)[0].astype(np.int32))
    assert np.array_equal(v.view(np.uint64), ref[keep].view(np.uint64))
    lines, exact = info.stats()
    assert lines == n and exact > 500  # ties and long tokens went through the exact conversion
    A32 = sp.create_COO_from_MM(info, np.float32)
    with np.errstate(over="ignore"):
        assert np.array_equal(A32.val.numpy().view(np.uint32), ref[keep].astype(np.float32).view(np.uint32))


def test_mm_lenient_layout_equals_reference_layout(ctx):
    """Tabs, runs of blanks, CRLF, blank and comment lines in the body, and
    a last line without '\\n' read as the single-space LF file does."""
    text = random_mm(9, "real", "symmetric", n=40, nent=200)
    lines = text.split(b"\n")[:-1]
    head, body = lines[:3], lines[3:]
    messy = []
    for i, ln in enumerate(body):
        a, b, c = ln.split(b" ")
        if i % 3 == 0:
            ln = b"  " + a + b"\t" + b + b"   " + c + b"\r"
        elif i % 3 == 1:
            ln = a + b" " + b + b"\t" + c
        messy.append(ln)
        if i % 17 == 0:
            messy.append(b"")
        if i % 23 == 0:
            messy.append(b"% mid-body comment")
    messy_text = b"\r\n".join(head[:1]) + b"\r\n" + b"\n".join(head[1:] + messy)  # no final '\n'
    for dtype in (np.float32, np.float64):
        want = oracle.mm_coo(text, dtype)
        info = sp.read_matrix_MM_files(messy_text, ctx)
        A = sp.create_COO_from_MM(info, dtype)
        gr, gc, gv = coo_host(A)
        assert np.array_equal(gr, want[1]) and np.array_equal(gc, want[2])
        assert np.array_equal(gv.view(np.uint8), want[3].view(np.uint8))


def test_mm_device_text_and_mmread(ctx, tmp_path):
    text = random_mm(4, "real", "general", n=120, nent=2000)
    d = sp.DeviceArray.from_host(np.frombuffer(text, np.uint8), ctx=ctx)
    info = sp.read_matrix_MM_files(d, ctx)
    A = sp.create_COO_from_MM(info, np.float64)
    want = oracle.mm_coo(text, np.float64)
    assert np.array_equal(A.val.numpy(), want[3])
    path = tmp_path / "a.mtx"
    path.write_bytes(text)
    B = sp.mmread(str(path), np.float64, ctx)
    assert np.array_equal(B.row.numpy(), want[1]) and np.array_equal(B.col.numpy(), want[2])
    # mmread -> coo_csr, the reference's test flow (seq/tests/main32.js:24-30)
    csr = sp.coo_csr(B)
    rp, col, val = oracle.coo_csr(B.N, want[1], want[2], want[3])
    assert np.array_equal(csr.row.numpy(), rp) and np.array_equal(csr.col.numpy(), col)
    assert np.array_equal(csr.val.numpy(), val)


def test_mm_all_zero_and_empty(ctx):
    t = b"%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 0\n2 3 -0.0\n"
    info = sp.read_matrix_MM_files(t, ctx)
    assert info.nnz == 2
    A = sp.create_COO_from_MM(info, np.float32)
    assert A.row.numpy().tolist() == [0, 0] and A.val.numpy().tolist() == [0, 0]
    assert np.array_equal(A.val.numpy(), oracle.mm_coo(t)[3])
    t = b"%%MatrixMarket matrix coordinate real general\n% nothing\n4 4 0\n"
    info = sp.read_matrix_MM_files(t, ctx)
    assert info.nnz == 0 and info.lower_nnz == 4
    L = sp.create_LCOO_from_MM(info, np.float64)
    assert L.row.numpy().tolist() == [0, 1, 2, 3] and L.val.numpy().tolist() == [1, 1, 1, 1]


@pytest.mark.parametrize("text,needle", [
    (b"%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n", "index outside"),
    (b"%%MatrixMarket matrix coordinate real general\n2 2 1\n1 0 1.0\n", "index outside"),
    (b"%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 abc\n", "value"),
    (b"%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1\n", "value"),
    (b"%%MatrixMarket matrix coordinate real general\n2 2 1\n1.5 1 2\n", "row index"),
    (b"%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 2\n", "size line says 2"),
    (b"%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 2\n2 2 3\n", "size line says 1"),
    (b"%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", "coordinate"),
    (b"%%MatrixMarket matrix coordinate real symmetric\n2 3 1\n1 1 2\n", "square"),
    (b"not a banner\n1 1 1\n1 1 1\n", "banner"),
    (b"%%MatrixMarket matrix coordinate real general\n% no size line\n", "size line"),
])
def test_mm_errors(ctx, text, needle):
    with pytest.raises(SableError) as e:
        sp.read_matrix_MM_files(text, ctx)
    assert needle in str(e.value), str(e.value)
