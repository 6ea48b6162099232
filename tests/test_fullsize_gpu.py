"""Parity at the benchmarked workloads' full sizes, each test a few seconds
to about twenty.

* FP64 CSR on the Zipf(s = 1.5) matrix of 4,194,304 rows (~206M nnz) and
  on the Zipf(s = 1.73) matrix of the same N (~67M nnz): the column-sorted
  tile product (the headline schedule) and the autotuned plan against the
  oracle's reference-order CSR product, on a sample of rows at the start,
  middle and end, within the CSR bound
  |y - y_ref| <= 2 k u / (1 - k u) * sum_j |a_ij x_j| (k = row length + 1).
* The 7-diagonal band of 16,777,216 rows: csr_dia and csr_ell bit-exact
  against the oracle's conversions, and the DIA and ELL products bit-exact
  against the oracle's reference-order products, on every row.
* The lower solves of the 3D Laplacian 128^3 and of the random 1,048,576-row
  lower matrix: the analysis's default and tuned level solve bit-exact
  against the oracle's level solve, the sync-free solve against the
  oracle's sync-free solve, and the relative residual of both <= 1e-12; the
  fused FP32 + FP64-refinement solve the same on every schedule, exact on
  the leading rows and within 1e-12.
* The 2,048-matrix batch: the batched tile product of every matrix within
  its CSR bound of the oracle, and the batched sync-free solves of all
  2,048 diagonally dominant lower triangles bit-exact against the oracle.
"""
import numpy as np
import pytest

import oracle
import sciwasm.sparse_amd as sp
from sciwasm.sparse_amd import gen, DeviceArray

pytestmark = pytest.mark.gpu

N_HEAD = 4194304
SEED = 20261203


def csr_bound(rp, col, val, x, rows):
    """The CSR bound for the rows [r0, r1) of each range in `rows`."""
    u = np.finfo(np.float64).eps / 2
    out = []
    for r0, r1 in rows:
        a, b = int(rp[r0]), int(rp[r1])
        lens = np.diff(rp[r0:r1 + 1]).astype(np.float64) + 1
        ab = np.abs(val[a:b] * x[col[a:b]])
        rid = np.repeat(np.arange(r1 - r0), np.diff(rp[r0:r1 + 1]))
        s = np.bincount(rid, weights=ab, minlength=r1 - r0)
        out.append(2 * lens * u / (1 - lens * u) * s + 1e-300)
    return out


def sample_rows(N, n=65536):
    return [(0, n), (N // 2 - n // 2, N // 2 + n // 2), (N - n, N)]


def check_rows(y, rp, col, val, x, rows):
    for (r0, r1), bnd in zip(rows, csr_bound(rp, col, val, x, rows)):
        a = int(rp[r0])
        lrp = (rp[r0:r1 + 1] - a).astype(np.int32)
        yref = oracle.spmv_csr(r1 - r0, lrp, col[a:int(rp[r1])], val[a:int(rp[r1])], x, np.zeros(r1 - r0))
        d = np.abs(y[r0:r1] - yref)
        assert np.all(d <= bnd), (r0, r1, float(np.max(d / bnd)))


@pytest.mark.parametrize("s", [1.5, 1.73])
def test_zipf_fullsize_tiles_and_autotuned(ctx, s):
    rp, col, val = gen.zipf_csr(N_HEAD, s, 4096, seed=SEED, dtype=np.float64)
    x = np.random.default_rng(SEED + 1).uniform(-1.0, 1.0, N_HEAD)
    print("zipf s=%.2f nnz=%d" % (s, len(col)), flush=True)
    A = sp.csr_from_host(N_HEAD, rp, col, val)
    xv = sp.vec_from_host(x)
    rows = sample_rows(N_HEAD)
    tl = sp.CsrTiles(A, 16384, 8192 if s == 1.5 else 16384)
    yv = sp.create_vec(N_HEAD, 0.0, np.float64)
    tl.run(xv.vec, yv.vec, beta=0.0)
    check_rows(yv.array(), rp, col, val, x, rows)
    del tl
    auto = sp.CsrAuto(A)
    print("autotuned", auto.info(), flush=True)
    yv = sp.create_vec(N_HEAD, 0.0, np.float64)
    auto.run(xv.vec, yv.vec, beta=0.0)
    check_rows(yv.array(), rp, col, val, x, rows)


def test_band_fullsize_dia_ell_bit_exact(ctx):
    N = 1 << 24
    offs = [-4096, -64, -1, 0, 1, 64, 4096]
    rp, col, val = gen.band_csr(N, offs, seed=5)
    x = np.random.default_rng(2).uniform(-1, 1, N)
    A = sp.csr_from_host(N, rp, col, val)
    xv = sp.vec_from_host(x)
    D = sp.csr_dia(A)
    nd, stride, eoff, edata = oracle.csr_dia(N, rp, col, val)
    assert D.ndiags == nd and D.stride == stride
    assert np.array_equal(D.offset.numpy(), eoff) and np.array_equal(D.data.numpy(), edata)
    yv = sp.create_vec(N, 0.0, np.float64)
    sp.spmv(D, xv, yv, beta=0.0)
    assert np.array_equal(yv.array(), oracle.spmv_dia(N, stride, eoff, edata, x, np.zeros(N)))
    del D, edata
    E = sp.csr_ell(A)
    nc, eidx, edat = oracle.csr_ell(N, rp, col, val)
    assert E.ncols == nc
    assert np.array_equal(E.indices.numpy(), eidx) and np.array_equal(E.data.numpy(), edat)
    yv = sp.create_vec(N, 0.0, np.float64)
    sp.spmv(E, xv, yv, beta=0.0)
    assert np.array_equal(yv.array(), oracle.spmv_ell(N, nc, eidx, edat, x, np.zeros(N)))


def level_solution(N, rp, c, v, b):
    nl, lp, perm, order, _ = oracle.level_sets(N, rp, c)
    nrp, nc_, nv = oracle.level_reorder(N, rp, c, v, perm, order)
    yl = oracle.spts_level(N, nrp, nc_, nv, b[perm])
    x = np.empty_like(yl)
    x[perm] = yl
    return x


@pytest.mark.parametrize("name", ["lap3d_128", "random_lower_1M"])
def test_spts_fullsize_bit_exact(ctx, name):
    rp, c, v = gen.laplacian_3d_lower(128) if name == "lap3d_128" else gen.random_lower(1 << 20, 32, seed=7)
    N = len(rp) - 1
    b = np.random.default_rng(8).uniform(-1, 1, N)
    an = sp.SptsAnalysis(sp.csr_from_host(N, rp, c, v))
    bv = sp.vec_from_host(b)
    want_l = level_solution(N, rp, c, v, b)
    want_s = oracle.spts_syncfree(N, rp, c, v, b)
    for label in ("default", "tuned"):
        if label == "tuned":
            print(name, an.tune(), flush=True)
        x = sp.create_vec(N, 0.0, np.float64)
        an.solve_level(bv, x)
        an.status()
        assert np.array_equal(x.array(), want_l), (name, label)
        x = sp.create_vec(N, 0.0, np.float64)
        an.solve_syncfree(bv, x)
        an.status()
        assert np.array_equal(x.array(), want_s), (name, label)
    assert oracle.rel_residual(N, rp, c, v, want_l, b) <= 1e-12
    assert oracle.rel_residual(N, rp, c, v, want_s, b) <= 1e-12


@pytest.mark.parametrize("name", ["lap3d_128", "random_lower_1M"])
def test_spts_fullsize_fused_refinement(ctx, name):
    """The FP32 solve with one FP64 refinement step fused into its sweep, at
    full size: the same x on the row blocks, the level graph at 8 and 64
    lanes and the tuned schedule, the leading 20,000 rows (a lower matrix's
    leading block depends on nothing after it) equal to the exact host
    computation of the steps, and a relative residual <= 1e-12."""
    from test_spts_gpu import host_fused_refine
    rp, c, v = gen.laplacian_3d_lower(128) if name == "lap3d_128" else gen.random_lower(1 << 20, 32, seed=7)
    N = len(rp) - 1
    b = np.random.default_rng(9).uniform(-1, 1, N)
    an = sp.SptsAnalysis(sp.csr_from_host(N, rp, c, v))
    bv = sp.vec_from_host(b)
    xs = {}
    for label in ("blocks", "graph8", "graph64", "tuned"):
        if label.startswith("graph"):
            an.set_lanes(int(label[5:]))
            an.set_schedule("graph")
        elif label == "tuned":
            an.tune()
        x = sp.create_vec(N, 0.0, np.float64)
        an.solve_refined(bv, x, iters=1, inner="fused")
        an.status()
        xs[label] = x.array()
    for label, xa in xs.items():
        assert np.array_equal(xa, xs["blocks"]), (name, label)
    K = 20000
    nk = int(rp[K])
    want = host_fused_refine(K, rp[:K + 1], c[:nk], v[:nk], b[:K])
    assert np.array_equal(xs["blocks"][:K], want), name
    assert oracle.rel_residual(N, rp, c, v, xs["blocks"], b) <= 1e-12


def test_batch_fullsize_tiles_and_solves(ctx):
    ns, rs, rp, col, val = gen.batch_csr_concat(2048, seed=SEED + 4)
    rng = np.random.default_rng(3)
    mats, xs, ys, X, host = [], [], [], [], []
    for k in range(len(ns)):
        a, b = rs[k], rs[k + 1]
        lrp = (rp[a:b + 1] - rp[a]).astype(np.int32)
        n = int(ns[k])
        lc, lv = col[rp[a]:rp[b]], val[rp[a]:rp[b]]
        host.append((n, lrp, lc, lv))
        mats.append(sp.csr_from_host(n, lrp, lc, lv))
        x = rng.uniform(-1, 1, n)
        X.append(x)
        xs.append(DeviceArray.from_host(x))
        ys.append(DeviceArray.zeros(n, np.float64))
    plans = sp.batch_tile_plans(mats)
    sp.batch_spmv_tiles(plans, xs, ys)
    ctx.sync()
    for k, (n, lrp, lc, lv) in enumerate(host):
        bnd = csr_bound(lrp, lc, lv, X[k], [(0, n)])[0]
        yref = oracle.spmv_csr(n, lrp, lc, lv, X[k], np.zeros(n))
        assert np.all(np.abs(ys[k].numpy() - yref) <= bnd), k
    del plans, mats, xs, ys
    Ls, bs, xs, want = [], [], [], []
    for n, lrp, lc, lv in host:
        L = sp.lower_dominant(lrp, lc, lv)
        b = rng.uniform(-1, 1, n)
        Ls.append(sp.csr_from_host(n, *L))
        bs.append(DeviceArray.from_host(b))
        xs.append(DeviceArray(n, np.float64))
        want.append(oracle.spts_syncfree(n, L[0], L[1], L[2], b))
    sp.batch_spts(Ls, bs, xs, lanes=4)
    for k in range(len(Ls)):
        assert np.array_equal(xs[k].numpy(), want[k]), k
