"""MatrixMarket ingest, CPU side.

* The oracle (oracle.mm_coo, a restatement of seq/src/sparse32.js:1889-2172)
  is pinned by hand-worked files and by scipy.io.mmread, an independent
  reader: the full COO must equal scipy's matrix with explicit zeros
  dropped, and the lower COO must equal tril(A, -1) + I.
* The device parser's number conversion (sciwasm/sparse_amd/csrc/
  mm_parse.h) is compiled for the host with g++ and checked bit for bit
  against Python's float(), which rounds correctly like JavaScript's Number()
  and C's strtod: repr strings, long and short digit strings, exact
  midpoints between doubles (ties), subnormals, overflow, and the literal
  forms Number() accepts or rejects.
* pow5_table.inc is regenerated and compared.
"""
import ctypes
import io
import os
import random
import struct
import subprocess
from decimal import Decimal, getcontext

import numpy as np
import pytest
import scipy.io
import scipy.sparse

import oracle
from sciwasm.sparse_amd import gen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "sciwasm", "sparse_amd", "csrc")


# ------------------------------------------------------------------ oracle
HAND = b"""%%MatrixMarket matrix coordinate real symmetric
% a comment
3 3 4
1 1 2.5
2 1 -1
3 2 0
3 3 4e0
"""


def test_oracle_hand_symmetric():
    info, r, c, v = oracle.mm_coo(HAND)
    # (3,2) is an explicit zero and is dropped; (2,1) is mirrored to (1,2)
    assert info == dict(nrows=3, ncols=3, nentries=4, nnz=4, pattern=False, symmetric=True)
    assert r.tolist() == [0, 0, 1, 2] and c.tolist() == [0, 1, 0, 2]
    assert v.tolist() == [2.5, -1.0, -1.0, 4.0] and v.dtype == np.float32


def test_oracle_hand_lower():
    info, r, c, v = oracle.mm_coo(HAND, np.float64, lower=True)
    # strictly lower nonzeros (2,1) plus a unit diagonal on all 3 rows
    assert info["nnz"] == 4
    assert r.tolist() == [0, 1, 1, 2] and c.tolist() == [0, 0, 1, 2]
    assert v.tolist() == [1.0, -1.0, 1.0, 1.0] and v.dtype == np.float64


def test_oracle_hand_pattern_symmetric():
    t = b"%%MatrixMarket matrix coordinate pattern symmetric\n3 3 3\n1 1\n2 1\n3 2\n"
    info, r, c, v = oracle.mm_coo(t)
    assert info["nnz"] == 5  # 1 + 2 + 2
    assert list(zip(r.tolist(), c.tolist())) == [(0, 0), (0, 1), (1, 0), (1, 2), (2, 1)]
    assert (v == 1).all()


def test_oracle_all_zero_values_fall_back_to_nentries():
    # read_matrix_MM_files: nnz == 0 -> nnz = nentries; nothing is written,
    # so the entries stay as allocated (zero)
    t = b"%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 0\n2 2 0.0\n"
    info, r, c, v = oracle.mm_coo(t)
    assert info["nnz"] == 2 and r.tolist() == [0, 0] and c.tolist() == [0, 0] and v.tolist() == [0, 0]


def test_oracle_unterminated_last_line_is_dropped():
    # the loader keeps the text after the last '\n' back (:2470-2485)
    t = b"%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 3\n2 2 4"
    info, r, c, v = oracle.mm_coo(t)
    assert info["nentries"] == 2 and info["nnz"] == 1 and v.tolist() == [3.0]


def test_oracle_general_lower_keeps_only_below_diagonal():
    t = b"%%MatrixMarket matrix coordinate real general\n3 3 4\n1 2 5\n2 1 6\n3 1 0\n3 3 7\n"
    info, r, c, v = oracle.mm_coo(t, lower=True)
    assert list(zip(r.tolist(), c.tolist(), v.tolist())) == [(0, 0, 1), (1, 0, 6), (1, 1, 1), (2, 2, 1)]


def random_mm(seed, field, symmetry, n=60, m=None, nent=300, style="repr"):
    rng = np.random.default_rng(seed)
    m = n if m is None else m
    keys = set()
    rows, cols = [], []
    while len(rows) < nent:
        r, c = int(rng.integers(1, n + 1)), int(rng.integers(1, m + 1))
        if symmetry == "symmetric" and r < c:
            r, c = c, r
        if (r, c) in keys:
            continue
        keys.add((r, c))
        rows.append(r)
        cols.append(c)
    pool = gen.value_pool(64, seed=seed, style=style) + [b"0", b"0.0", b"-0"]
    idx = rng.integers(0, len(pool), nent)
    if field == "pattern":
        return gen.mm_text(n, m, rows, cols, field=field, symmetry=symmetry)
    if field == "integer":
        pool = [b"%d" % k for k in range(-5, 6)]
        idx = rng.integers(0, len(pool), nent)
    return gen.mm_text(n, m, rows, cols, pool, idx, field=field, symmetry=symmetry)


def scipy_full(text):
    A = scipy.io.mmread(io.BytesIO(text))
    A = scipy.sparse.coo_matrix(A)
    keep = A.data != 0
    A = scipy.sparse.coo_matrix((A.data[keep], (A.row[keep], A.col[keep])), shape=A.shape)
    return A


@pytest.mark.parametrize("field,symmetry", [("real", "general"), ("real", "symmetric"), ("integer", "general"),
                                            ("pattern", "general"), ("pattern", "symmetric")])
@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_matches_scipy_reader(field, symmetry, seed):
    text = random_mm(seed, field, symmetry, n=50)
    info, r, c, v = oracle.mm_coo(text, np.float64)
    A = scipy_full(text)
    order = np.lexsort((A.col, A.row))
    assert info["nnz"] == A.nnz
    assert np.array_equal(r, A.row[order]) and np.array_equal(c, A.col[order])
    assert np.array_equal(v, A.data[order].astype(np.float64))
    # sorted by (row, col), as quick_sort_COO leaves it
    key = r.astype(np.int64) * 1000 + c
    assert (np.diff(key) > 0).all()
    # lower: tril(A, -1) + I
    info, r, c, v = oracle.mm_coo(text, np.float64, lower=True)
    L = (scipy.sparse.tril(A, -1) + scipy.sparse.eye(A.shape[0])).tocoo()
    order = np.lexsort((L.col, L.row))
    assert info["nnz"] == L.nnz
    assert np.array_equal(r, L.row[order]) and np.array_equal(c, L.col[order])
    assert np.array_equal(v, L.data[order])


def test_oracle_matches_scipy_reader_rectangular():
    text = random_mm(5, "real", "general", n=40, m=90)
    info, r, c, v = oracle.mm_coo(text, np.float64)
    A = scipy_full(text)
    order = np.lexsort((A.col, A.row))
    assert (info["nrows"], info["ncols"], info["nnz"]) == (40, 90, A.nnz)
    assert np.array_equal(r, A.row[order]) and np.array_equal(c, A.col[order])
    assert np.array_equal(v, A.data[order])


def test_oracle_values_round_like_float():
    toks = [b"0.1", b"1e-310", b"2.4703282292062328e-324", b"9007199254740993", b"-1.7976931348623157e308",
            b"123456789012345678901234567890", b"3.4028235677973366e38"]
    rows = np.arange(1, len(toks) + 1)
    text = gen.mm_text(len(toks), 1, rows, np.ones(len(toks), np.int64), toks, np.arange(len(toks)))
    _, _, _, v64 = oracle.mm_coo(text, np.float64)
    _, _, _, v32 = oracle.mm_coo(text, np.float32)
    ref = np.array([float(t) for t in toks])
    assert np.array_equal(v64.view(np.uint64), ref.view(np.uint64))
    with np.errstate(over="ignore"):
        assert np.array_equal(v32.view(np.uint32), ref.astype(np.float32).view(np.uint32))


# ------------------------------------------------------------------ parser, host build
@pytest.fixture(scope="module")
def host_parser(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("mmh") / "mm_parse_host.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", so,
                           os.path.join(ROOT, "tests", "native", "mm_parse_host.cpp")])
    return ctypes.CDLL(so)


def _pack(strs):
    text = b"".join(strs)
    off = np.zeros(len(strs) + 1, np.int64)
    off[1:] = np.cumsum([len(s) for s in strs])
    return text, off


def parse_host(lib, strs):
    text, off = _pack(strs)
    out = np.zeros(len(strs), np.uint64)
    path = np.zeros(len(strs), np.int32)
    lib.host_parse_decimals(ctypes.c_char_p(text), ctypes.c_void_p(off.ctypes.data), ctypes.c_int64(len(strs)),
                            ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(path.ctypes.data))
    return out, path


def bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def decimal_cases(seed, n_rand=20000, n_mid=3000):
    """Strings that stress every branch of the conversion."""
    rng = random.Random(seed)
    out = []
    for _ in range(n_rand):
        x = struct.unpack("<d", struct.pack("<Q", rng.getrandbits(64)))[0]
        if x != x or abs(x) == float("inf"):
            continue
        out.append(repr(x).encode())
        out.append(b"%.*e" % (rng.randint(0, 25), x))
    for _ in range(n_rand):
        nd = rng.randint(1, 30)
        dg = "".join(rng.choice("0123456789") for _ in range(nd))
        e = rng.randint(-360, 330)
        k = rng.randint(0, nd)
        out.append(("%se%d" % (dg, e)).encode())
        out.append(("%s%s.%sE%d" % ("-" if rng.random() < 0.5 else "", dg[:k], dg[k:], e)).encode())
    getcontext().prec = 2000
    for _ in range(n_mid):
        if rng.random() < 0.5:
            M, e = rng.randint(2 ** 52, 2 ** 53 - 1), rng.randint(-1074, 971)
        else:
            M, e = rng.randint(0, 2 ** 52 - 1), -1074
        mid = Decimal(2 * M + 1) * (Decimal(2) ** (e - 1))  # exact midpoint between two doubles
        out.append((format(mid, "f") if abs(mid.adjusted()) < 30 else format(mid, "e")).encode())
        for nd in (17, 19, 20, 25, 40):
            out.append(format(mid, ".%de" % (nd - 1)).encode())
    out += [b"0", b"-0", b"0.0e10", b".5", b"5.", b"+1", b"-Infinity", b"Infinity", b"1e400", b"1e-400",
            b"2.4703282292062327e-324", b"2.4703282292062328e-324", b"4.9e-324", b"1.7976931348623157e308",
            b"1.7976931348623158e308", b"1.7976931348623159e308", b"9007199254740993", b"9007199254740995",
            b"0.5", b"0.25", b"1.5", b"-2.5", b"0.125", b"3.0000000000000000000000001", b"1e23",
            b"2.2250738585072011e-308", b"2.2250738585072012e-308", b"123456789012345678901234567890e-10",
            b"0." + b"0" * 400 + b"1e400", b"1" + b"0" * 500 + b"e-500"]
    return out


def test_host_parser_matches_float(host_parser):
    strs = decimal_cases(7)
    out, path = parse_host(host_parser, strs)
    ref = np.array([bits(float(s)) for s in strs], np.uint64)
    bad = [(s, hex(int(o)), hex(int(r)), p) for s, o, r, p in zip(strs, out, ref, path) if p < 0 or o != r]
    assert not bad, bad[:10]
    assert (path == 1).sum() > 1000  # the exact path ran (ties, long digit strings)


def test_host_exact_path_alone(host_parser):
    """exact_bits() from candidates up to 3 units off must land on float()."""
    strs = [s for s in decimal_cases(11, 5000, 1000) if b"Infinity" not in s]
    text, off = _pack(strs)
    ref = np.array([bits(abs(float(s))) for s in strs], np.int64)
    rng = np.random.default_rng(3)
    cand = np.clip(ref + rng.integers(-3, 4, len(strs)), 0, 0x7FF0000000000000).astype(np.uint64)
    out = np.zeros(len(strs), np.uint64)
    fl = np.zeros(len(strs), np.int32)
    host_parser.host_exact_decimals(ctypes.c_char_p(text), ctypes.c_void_p(off.ctypes.data),
                                    ctypes.c_int64(len(strs)), ctypes.c_void_p(cand.ctypes.data),
                                    ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(fl.ctypes.data))
    sref = np.array([bits(float(s)) for s in strs], np.uint64)
    assert not fl.any()
    assert np.array_equal(out, sref)


def test_host_parser_rejects_what_number_rejects(host_parser):
    bad = [b"", b"e5", b".", b"1e", b"1e+", b"abc", b"1.2.3", b"--1", b"inf", b"nan", b"0x10", b"1,5", b"+", b"-.",
           b"Infinit", b"Infinityx"]
    _, path = parse_host(host_parser, [s + b" " for s in bad])
    assert (path == -1).all()


def test_host_line_parser(host_parser):
    lines = [b"1 2 3.5", b"  7\t8   -1e3 extra", b"", b"   ", b"% c", b"1", b"x 2 3", b"1 y 3", b"1 2 z", b"+4 5 6",
             b"1 2"]
    text, off = _pack([ln + b"\n" for ln in lines])
    n = len(lines)
    row = np.zeros(n, np.int64)
    col = np.zeros(n, np.int64)
    st = np.zeros(n, np.int32)
    host_parser.host_parse_lines(ctypes.c_char_p(text), ctypes.c_void_p(off.ctypes.data), ctypes.c_int64(n), 0,
                                 ctypes.c_void_p(row.ctypes.data), ctypes.c_void_p(col.ctypes.data),
                                 ctypes.c_void_p(st.ctypes.data))
    assert st.tolist() == [1, 1, 0, 0, 0, -2, -1, -2, -3, 1, -3]
    assert row[:2].tolist() == [1, 7] and col[:2].tolist() == [2, 8] and row[9] == 4


def test_pow5_table_is_generated(tmp_path):
    out = str(tmp_path / "pow5.inc")
    subprocess.check_call(["python3", os.path.join(CSRC, "gen_pow5.py"), out])
    with open(out) as a, open(os.path.join(CSRC, "pow5_table.inc")) as b:
        assert a.read() == b.read()
