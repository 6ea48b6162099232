"""Seeded synthetic workloads (host numpy) for the hot-path configurations.

* ``laplacian_2d(n)``: 5-point Laplacian on an n x n grid, natural order,
  diagonal 4, neighbours -1 (65,536 rows and 326,656 nonzeros for n=256).
* ``zipf_csr(N, s, kmax, seed)``: N x N, row lengths from the Zipf law
  P(k) ~ k^-s truncated to [1, kmax], per-row columns uniform without
  replacement (sorted), values N(0,1).
* ``band_csr(N, offsets, seed)``: banded matrix with the given diagonals,
  values N(0,1).
* ``laplacian_3d_lower(n)``: lower triangle, diagonal included, of the
  7-point Laplacian on an n^3 grid (diagonal 6, off-diagonals -1).
* ``random_lower(N, kmax, seed)``: row i has min(U{1..kmax}, i) distinct
  uniform columns < i, values U(-1,1)/row_len, diagonal 1 + sum|off|.
* ``batch_csr(count, ...)``: many independent random CSR matrices.

All return (row_ptr int32, col int32, val float64) unless noted; row
columns ascend and a lower matrix keeps its diagonal last.
"""
import concurrent.futures
import os

import numpy as np


def _workers():
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 4
    return max(1, min(16, n))


def laplacian_2d(n, dtype=np.float64):
    N = n * n
    i = np.arange(N, dtype=np.int64)
    gx, gy = i % n, i // n
    cols = []
    vals = []
    for off, valid, v in ((-n, gy > 0, -1.0), (-1, gx > 0, -1.0), (0, np.ones(N, bool), 4.0),
                          (1, gx < n - 1, -1.0), (n, gy < n - 1, -1.0)):
        cols.append(np.where(valid, i + off, -1))
        vals.append(np.full(N, v))
    C = np.stack(cols, 1)
    V = np.stack(vals, 1)
    mask = C >= 0
    row_ptr = np.zeros(N + 1, np.int64)
    row_ptr[1:] = np.cumsum(mask.sum(1))
    return row_ptr.astype(np.int32), C[mask].astype(np.int32), V[mask].astype(dtype)


def band_csr(N, offsets, seed=0, dtype=np.float64):
    offsets = np.sort(np.asarray(offsets, np.int64))
    i = np.arange(N, dtype=np.int64)
    counts = np.zeros(N, np.int64)
    for o in offsets:
        counts += ((i + o >= 0) & (i + o < N))
    row_ptr = np.zeros(N + 1, np.int64)
    row_ptr[1:] = np.cumsum(counts)
    nnz = int(row_ptr[-1])
    col = np.empty(nnz, np.int32)
    rank = np.zeros(N, np.int64)
    for o in offsets:
        v = (i + o >= 0) & (i + o < N)
        rows = i[v]
        col[row_ptr[rows] + rank[rows]] = (rows + o).astype(np.int32)
        rank[rows] += 1
    rng = np.random.default_rng(seed)
    val = rng.standard_normal(nnz).astype(dtype)
    return row_ptr.astype(np.int32), col, val


def zipf_lengths(N, s, kmax, rng):
    k = np.arange(1, kmax + 1, dtype=np.float64)
    p = k ** (-float(s))
    cdf = np.cumsum(p)
    cdf /= cdf[-1]
    u = rng.random(N)
    return (np.searchsorted(cdf, u, side="right") + 1).clip(1, kmax).astype(np.int64)


def _sorted_distinct_cols(lens, N, rng):
    """Per row, len distinct uniform columns in [0, N), ascending (N may be
    a per-row array)."""
    m = lens.astype(np.int64)
    tot = int(m.sum())
    nrow = len(m)
    E = rng.standard_exponential(tot + nrow)
    cs = np.cumsum(E)
    grp_end = np.cumsum(m + 1) - 1          # index of the (k+1)-th spacing of each row
    grp_start = grp_end - m                 # index of the first spacing
    base = np.where(grp_start > 0, cs[np.maximum(grp_start - 1, 0)], 0.0)
    total = cs[grp_end] - base
    keep = np.ones(tot + nrow, bool)
    keep[grp_end] = False
    rid = np.repeat(np.arange(nrow), m)
    u = (cs[keep] - base[rid]) / total[rid]
    Nr = N[rid] if isinstance(N, np.ndarray) else N
    cols = np.minimum((u * Nr).astype(np.int64), np.asarray(Nr) - 1)
    # repair the rare rows with a repeated column
    dup = np.zeros(tot, bool)
    if tot > 1:
        dup[1:] = (cols[1:] == cols[:-1]) & (rid[1:] == rid[:-1])
    if dup.any():
        row_start = np.zeros(nrow + 1, np.int64)
        row_start[1:] = np.cumsum(m)
        for r in np.unique(rid[dup]):
            a, b = row_start[r], row_start[r + 1]
            cols[a:b] = np.sort(rng.choice(int(N[r]) if isinstance(N, np.ndarray) else N, b - a, replace=False))
    return cols.astype(np.int32)


def zipf_csr(N, s=1.5, kmax=4096, seed=0, dtype=np.float64, chunk_rows=1 << 18):
    """Zipf(s) row lengths truncated to [1, kmax] (renormalised law)."""
    ss = np.random.SeedSequence(seed)
    rng0 = np.random.default_rng(ss.spawn(1)[0])
    lens = zipf_lengths(N, s, kmax, rng0)
    row_ptr = np.zeros(N + 1, np.int64)
    row_ptr[1:] = np.cumsum(lens)
    nnz = int(row_ptr[-1])
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, dtype)
    chunks = [(a, min(N, a + chunk_rows)) for a in range(0, N, chunk_rows)]
    seeds = ss.spawn(len(chunks) + 1)[1:]

    def work(args):
        (a, b), sd = args
        rng = np.random.default_rng(sd)
        lo, hi = row_ptr[a], row_ptr[b]
        col[lo:hi] = _sorted_distinct_cols(lens[a:b], N, rng)
        val[lo:hi] = rng.standard_normal(hi - lo).astype(dtype)

    with concurrent.futures.ThreadPoolExecutor(_workers()) as ex:
        list(ex.map(work, zip(chunks, seeds)))
    return row_ptr.astype(np.int32), col, val


def laplacian_3d_lower(n, dtype=np.float64):
    N = n ** 3
    i = np.arange(N, dtype=np.int64)
    gx, gy, gz = i % n, (i // n) % n, i // (n * n)
    C = np.stack([np.where(gz > 0, i - n * n, -1), np.where(gy > 0, i - n, -1),
                  np.where(gx > 0, i - 1, -1), i], 1)
    V = np.stack([np.full(N, -1.0)] * 3 + [np.full(N, 6.0)], 1)
    mask = C >= 0
    row_ptr = np.zeros(N + 1, np.int64)
    row_ptr[1:] = np.cumsum(mask.sum(1))
    return row_ptr.astype(np.int32), C[mask].astype(np.int32), V[mask].astype(dtype)


def random_lower(N, kmax=32, seed=0, dtype=np.float64):
    rng = np.random.default_rng(seed)
    want = rng.integers(1, kmax + 1, N)
    k = np.minimum(want, np.arange(N))           # off-diagonals in row i
    row_ptr = np.zeros(N + 1, np.int64)
    row_ptr[1:] = np.cumsum(k + 1)
    nnz = int(row_ptr[-1])
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    # distinct uniform columns < i: sample with replacement, repair repeats
    rid = np.repeat(np.arange(N), k)
    c = (rng.random(len(rid)) * rid).astype(np.int64)
    order = np.lexsort((c, rid))
    c = c[order]
    dup = np.zeros(len(c), bool)
    dup[1:] = (c[1:] == c[:-1]) & (rid[1:] == rid[:-1])
    starts = np.zeros(N + 1, np.int64)
    starts[1:] = np.cumsum(k)
    for r in np.unique(rid[dup]):
        a, b = starts[r], starts[r + 1]
        c[a:b] = np.sort(rng.choice(r, b - a, replace=False))
    offv = rng.uniform(-1.0, 1.0, len(rid)) / np.maximum(k[rid], 1)
    # place off-diagonals then the diagonal last
    pos_off = row_ptr[rid] + (np.arange(len(rid)) - starts[rid])
    col[pos_off] = c
    val[pos_off] = offv
    absum = np.bincount(rid, weights=np.abs(offv), minlength=N)
    dpos = row_ptr[1:] - 1
    col[dpos] = np.arange(N)
    val[dpos] = 1.0 + absum
    return row_ptr.astype(np.int32), col, val.astype(dtype)


def random_csr(N, nnz_per_row=5, seed=0, dtype=np.float64, duplicates=False, empty_rows=0.0):
    """Small random square CSR for tests (ascending columns; optional
    repeated (i,j) pairs and empty rows)."""
    rng = np.random.default_rng(seed)
    lens = rng.poisson(nnz_per_row, N).clip(0, N)
    if empty_rows > 0:
        lens[rng.random(N) < empty_rows] = 0
    rows, cols = [], []
    for i in range(N):
        if duplicates:
            c = np.sort(rng.integers(0, N, lens[i]))
        else:
            c = np.sort(rng.choice(N, lens[i], replace=False))
        rows.append(np.full(len(c), i))
        cols.append(c)
    r = np.concatenate(rows) if rows else np.zeros(0, np.int64)
    c = np.concatenate(cols) if cols else np.zeros(0, np.int64)
    row_ptr = np.zeros(N + 1, np.int64)
    row_ptr[1:] = np.cumsum(lens)
    val = rng.standard_normal(len(c)).astype(dtype)
    return row_ptr.astype(np.int32), c.astype(np.int32), val


def batch_csr(count, rows_lo=512, rows_hi=131072, lam=12, kmax=64, seed=0, dtype=np.float64):
    """Independent square CSR matrices: rows log-uniform in [rows_lo, rows_hi],
    nnz/row Poisson(lam) clipped to [1, kmax], uniform columns, N(0,1)."""
    ss = np.random.SeedSequence(seed)
    out = []
    rngs = [np.random.default_rng(s) for s in ss.spawn(count)]
    for rng in rngs:
        n = int(np.exp(rng.uniform(np.log(rows_lo), np.log(rows_hi))))
        lens = rng.poisson(lam, n).clip(1, min(kmax, n))
        row_ptr = np.zeros(n + 1, np.int64)
        row_ptr[1:] = np.cumsum(lens)
        col = _sorted_distinct_cols(lens, n, rng)
        val = rng.standard_normal(int(row_ptr[-1])).astype(dtype)
        out.append((n, row_ptr.astype(np.int32), col, val))
    return out


def csr_to_coo_rows(row_ptr):
    N = len(row_ptr) - 1
    return np.repeat(np.arange(N, dtype=np.int32), np.diff(row_ptr))


def batch_csr_concat(count, rows_lo=512, rows_hi=131072, lam=12, kmax=64, seed=0, dtype=np.float64,
                     chunk_rows=1 << 20):
    """The batch_csr law for many matrices at once, concatenated: returns
    (ns, row_start, row_ptr, col, val) where matrix k owns global rows
    [row_start[k], row_start[k+1]), row_ptr is global (int64) and col holds
    local column indices."""
    ss = np.random.SeedSequence(seed)
    rng = np.random.default_rng(ss.spawn(1)[0])
    ns = np.exp(rng.uniform(np.log(rows_lo), np.log(rows_hi), count)).astype(np.int64)
    row_start = np.zeros(count + 1, np.int64)
    row_start[1:] = np.cumsum(ns)
    R = int(row_start[-1])
    Nrow = np.repeat(ns, ns)
    lens = np.minimum(np.maximum(rng.poisson(lam, R), 1), np.minimum(kmax, Nrow))
    row_ptr = np.zeros(R + 1, np.int64)
    row_ptr[1:] = np.cumsum(lens)
    nnz = int(row_ptr[-1])
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, dtype)
    chunks = [(a, min(R, a + chunk_rows)) for a in range(0, R, chunk_rows)]
    seeds = ss.spawn(len(chunks) + 1)[1:]

    def work(args):
        (a, b), sd = args
        r = np.random.default_rng(sd)
        lo, hi = row_ptr[a], row_ptr[b]
        col[lo:hi] = _sorted_distinct_cols(lens[a:b], Nrow[a:b], r)
        val[lo:hi] = r.standard_normal(hi - lo).astype(dtype)

    with concurrent.futures.ThreadPoolExecutor(_workers()) as ex:
        list(ex.map(work, zip(chunks, seeds)))
    return ns, row_start, row_ptr, col, val


def concat_matrix(k, ns, row_start, row_ptr, col, val):
    """Matrix k of a batch_csr_concat result as (n, row_ptr, col, val)."""
    a, b = row_start[k], row_start[k + 1]
    rp = (row_ptr[a:b + 1] - row_ptr[a]).astype(np.int32)
    return int(ns[k]), rp, col[row_ptr[a]:row_ptr[b]], val[row_ptr[a]:row_ptr[b]]


def _digits_into(out, pos, x, width):
    """Writes the decimal digits of x (>= 0, int64) at out[pos:pos+width]."""
    for k in range(int(width.max()) if len(width) else 0):
        m = k < width
        p10 = np.power(10, np.maximum(width - 1 - k, 0), dtype=np.int64)
        out[pos[m] + k] = (48 + (x[m] // p10[m]) % 10).astype(np.uint8)


def _ndigits(x):
    w = np.ones(len(x), np.int64)
    t = x.copy()
    while True:
        t //= 10
        nz = t > 0
        if not nz.any():
            return w
        w += nz


def mm_text(nrows, ncols, row, col, val_tokens=None, val_idx=None, field="real", symmetry="general",
            comment=b"% generated\n"):
    """Coordinate MatrixMarket text, built with numpy (no per-entry Python).

    row, col: 1-based indices.  Values are tokens from the pool val_tokens
    (bytes) chosen by val_idx; pattern files pass None.  Lines are
    "row col value\\n" with single spaces."""
    row = np.asarray(row, np.int64)
    col = np.asarray(col, np.int64)
    n = len(row)
    head = b"%%%%MatrixMarket matrix coordinate %s %s\n" % (field.encode(), symmetry.encode())
    head += comment + b"%d %d %d\n" % (nrows, ncols, n)
    lr, lc = _ndigits(row), _ndigits(col)
    lens = lr + 1 + lc + 1
    if val_tokens is not None:
        pool_len = np.array([len(t) for t in val_tokens], np.int64)
        maxl = int(pool_len.max())
        pool = np.full((len(val_tokens), maxl), 32, np.uint8)
        for i, t in enumerate(val_tokens):
            pool[i, :len(t)] = np.frombuffer(t, np.uint8)
        val_idx = np.asarray(val_idx, np.int64)
        lv = pool_len[val_idx]
        lens = lens + 1 + lv
    offs = np.zeros(n + 1, np.int64)
    np.cumsum(lens, out=offs[1:])
    offs += len(head)
    out = np.empty(int(offs[-1]), np.uint8)
    out[:len(head)] = np.frombuffer(head, np.uint8)
    base = offs[:-1]
    _digits_into(out, base, row, lr)
    out[base + lr] = 32
    _digits_into(out, base + lr + 1, col, lc)
    p = base + lr + 1 + lc
    if val_tokens is not None:
        out[p] = 32
        for k in range(maxl):
            m = k < lv
            out[p[m] + 1 + k] = pool[val_idx[m], k]
        p = p + 1 + lv
    out[p] = 10
    return out.tobytes()


def value_pool(count, seed=0, style="repr"):
    """Distinct decimal tokens of random doubles: 'repr' (shortest round
    trip) or 'e16' ('%.16e', 17 significant digits)."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(count) * np.exp(rng.uniform(-8, 8, count))
    if style == "repr":
        return [repr(float(x)).encode() for x in v]
    return [(b"%.16e" % float(x)) for x in v]
