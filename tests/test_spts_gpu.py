"""GPU lower triangular solve against the oracle.

* Level schedule (level_ptr, permutation, order, reordered CSR): bit-exact.
* Every schedule and lane count keeps the reference's per-row operation
  order (lanes form the products of a row in parallel and add them in
  stored order), so the level-set solve is bit-identical to
  spts_csr/spts_level_csr and the sync-free solve to spts_csr_sync_free
  (both restated in the oracle).
* FP32 solve + FP64 refinement must reach the acceptance bound of the FP64
  path, relative residual ||b - L x||_inf / (||L||_inf ||x||_inf +
  ||b||_inf) <= 1e-12 against the long-double oracle (FP32 solves:
  <= 1e-5)."""
import numpy as np
import pytest

import oracle
import sciwasm.sparse_amd as sp
from sciwasm.sparse_amd import gen, SableError

pytestmark = pytest.mark.gpu

MATS = {
    "lap3d_16": lambda: gen.laplacian_3d_lower(16),
    "random_lower_20k": lambda: gen.random_lower(20000, 32, seed=3),
    "random_lower_sparse": lambda: gen.random_lower(5000, 2, seed=4),
}


def host_level_solution(N, rp, c, v, b):
    nl, lp, perm, order, lvl = oracle.level_sets(N, rp, c)
    nrp, nc_, nv = oracle.level_reorder(N, rp, c, v, perm, order)
    yl = oracle.spts_level(N, nrp, nc_, nv, b[perm])
    x = np.empty_like(yl)
    x[perm] = yl
    return x


@pytest.mark.parametrize("name", list(MATS))
def test_schedule_bit_exact(ctx, name):
    rp, c, v = MATS[name]()
    N = len(rp) - 1
    an = sp.SptsAnalysis(sp.csr_from_host(N, rp, c, v))
    lp, perm, order, nrp, ncol, nval = an.schedule()
    enl, elp, eperm, eorder, _ = oracle.level_sets(N, rp, c)
    assert an.nlevels == enl
    assert np.array_equal(lp, elp) and np.array_equal(perm, eperm) and np.array_equal(order, eorder)
    erp, ec, ev = oracle.level_reorder(N, rp, c, v, eperm, eorder)
    assert np.array_equal(nrp, erp) and np.array_equal(ncol, ec) and np.array_equal(nval, ev)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("name", list(MATS))
def test_lanes1_bit_exact(ctx, dtype, name):
    rp, c, v = MATS[name]()
    v = v.astype(dtype)
    N = len(rp) - 1
    b = np.random.default_rng(1).uniform(-1, 1, N).astype(dtype)
    L = sp.csr_from_host(N, rp, c, v)
    an = sp.SptsAnalysis(L, lanes=1)
    bv = sp.vec_from_host(b)
    xl = sp.create_vec(N, 0.0, dtype)
    an.solve_level(bv, xl)
    assert np.array_equal(xl.array(), host_level_solution(N, rp, c, v, b))
    xs = sp.create_vec(N, 0.0, dtype)
    an.solve_syncfree(bv, xs)
    an.status()
    assert np.array_equal(xs.array(), oracle.spts_syncfree(N, rp, c, v, b))
    xr = sp.create_vec(N, 0.0, dtype)
    sp.spts_syncfree(L, bv, xr, lanes=1)
    assert np.array_equal(xr.array(), xs.array())


@pytest.mark.parametrize("lanes", [4, 8, 16, 64])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("name", list(MATS))
def test_multilane_bit_exact(ctx, lanes, dtype, name):
    """More lanes per row in the level graph and the natural-order sync-free
    solve: the products of a row are formed in parallel and added in stored
    order, so the results are the reference solves' bit for bit."""
    rp, c, v = MATS[name]()
    v = v.astype(dtype)
    N = len(rp) - 1
    b = np.random.default_rng(2).uniform(-1, 1, N).astype(dtype)
    L = sp.csr_from_host(N, rp, c, v)
    an = sp.SptsAnalysis(L, lanes=lanes)
    an.set_schedule("graph")
    bv = sp.vec_from_host(b)
    want_l = host_level_solution(N, rp, c, v, b)
    want_s = oracle.spts_syncfree(N, rp, c, v, b)
    for _ in range(2):
        xl = sp.create_vec(N, 0.0, dtype)
        an.solve_level(bv, xl)
        xs = sp.create_vec(N, 0.0, dtype)
        an.solve_syncfree(bv, xs)
        an.status()
        assert np.array_equal(xl.array().view(np.uint8), want_l.view(np.uint8))
        assert np.array_equal(xs.array().view(np.uint8), want_s.view(np.uint8))
    xr = sp.create_vec(N, 0.0, dtype)
    sp.spts_syncfree(L, bv, xr, lanes=lanes)
    assert np.array_equal(xr.array().view(np.uint8), want_s.view(np.uint8))


@pytest.mark.parametrize("lanes", [1, 4, 16, 64])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("name", list(MATS))
def test_combined_equals_level(ctx, lanes, dtype, name):
    """The combined solve (sync-free over the level-ordered matrix) uses the
    level solve's arithmetic: bit-identical to it and to the reference-order
    level solve of the oracle at every lane count."""
    rp, c, v = MATS[name]()
    v = v.astype(dtype)
    N = len(rp) - 1
    b = np.random.default_rng(9).uniform(-1, 1, N).astype(dtype)
    an = sp.SptsAnalysis(sp.csr_from_host(N, rp, c, v), lanes=lanes)
    bv = sp.vec_from_host(b)
    xl = sp.create_vec(N, 0.0, dtype)
    an.set_schedule("graph")  # the per-level launches, with the analysis's lanes
    an.solve_level(bv, xl)
    for _ in range(2):
        xc = sp.create_vec(N, 0.0, dtype)
        an.solve_combined(bv, xc)
        an.status()
        assert np.array_equal(xc.array(), xl.array())
    assert np.array_equal(xc.array(), host_level_solution(N, rp, c, v, b))


def test_graph_replays_match_eager(ctx):
    rp, c, v = gen.random_lower(30000, 16, seed=5)
    N = len(rp) - 1
    an = sp.SptsAnalysis(sp.csr_from_host(N, rp, c, v), lanes=1)
    an.set_schedule("graph")
    rng = np.random.default_rng(6)
    for k in range(4):
        b = rng.uniform(-1, 1, N)
        x = sp.create_vec(N, 0.0, np.float64)
        an.solve_level(sp.vec_from_host(b), x)
        # eager work between the replays
        sp.spmv(sp.csr_from_host(N, rp, c, v), sp.vec_from_host(b), sp.create_vec(N, 1.0, np.float64))
        assert np.array_equal(x.array(), host_level_solution(N, rp, c, v, b)), k


def test_reference_shaped_level_solve(ctx):
    rp, c, v = gen.laplacian_3d_lower(12)
    N = len(rp) - 1
    A = sp.CSR_create_level_sets_with_reorder(sp.csr_from_host(N, rp, c, v), lanes=1)
    enl, elp, eperm, eorder, _ = oracle.level_sets(N, rp, c)
    assert A.nlevels == enl and np.array_equal(A.level, elp)
    b = np.random.default_rng(7).uniform(-1, 1, N)
    bl = sp.vec_from_host(b[eperm])
    yl = sp.create_vec(N, 0.0, np.float64)
    sp.spts_level_csr(A, bl, yl, lanes=1)
    erp, ec, ev = oracle.level_reorder(N, rp, c, v, eperm, eorder)
    assert np.array_equal(yl.array(), oracle.spts_level(N, erp, ec, ev, b[eperm]))


@pytest.mark.parametrize("inner", ["syncfree", "level"])
@pytest.mark.parametrize("name", ["lap3d_16", "random_lower_20k"])
def test_fp32_refinement_reaches_1e12(ctx, inner, name):
    rp, c, v = MATS[name]()
    N = len(rp) - 1
    b = np.random.default_rng(8).uniform(-1, 1, N)
    an = sp.SptsAnalysis(sp.csr_from_host(N, rp, c, v))
    x = sp.create_vec(N, 0.0, np.float64)
    an.solve_refined(sp.vec_from_host(b), x, iters=0, inner=inner)
    r0 = oracle.rel_residual(N, rp, c, v, x.array(), b)
    assert r0 < 1e-5
    an.solve_refined(sp.vec_from_host(b), x, iters=2, inner=inner)
    an.status()
    r2 = oracle.rel_residual(N, rp, c, v, x.array(), b)
    assert r2 <= 1e-12, (r0, r2)


def test_fp32_solve_residual(ctx):
    rp, c, v = gen.random_lower(20000, 32, seed=9, dtype=np.float32)
    N = len(rp) - 1
    b = np.random.default_rng(10).uniform(-1, 1, N).astype(np.float32)
    an = sp.SptsAnalysis(sp.csr_from_host(N, rp, c, v), lanes=8)
    x = sp.create_vec(N, 0.0, np.float32)
    an.solve_syncfree(sp.vec_from_host(b), x)
    assert oracle.rel_residual(N, rp, c, v.astype(np.float64), x.array().astype(np.float64),
                               b.astype(np.float64)) <= 1e-5


def test_malformed_and_singular(ctx):
    rp = np.array([0, 1, 3, 4], np.int32)
    c = np.array([0, 1, 0, 2], np.int32)        # row 1: diagonal not last
    L = sp.csr_from_host(3, rp, c, np.ones(4))
    with pytest.raises(SableError) as e:
        sp.SptsAnalysis(L)
    assert e.value.name == "SABLE_EMALFORMED"
    with pytest.raises(SableError) as e:
        sp.spts_syncfree(L, sp.create_vec(3, 1.0, np.float64), sp.create_vec(3, 0.0, np.float64))
    assert e.value.name == "SABLE_EMALFORMED"
    rp = np.array([0, 1, 3], np.int32)
    c = np.array([0, 0, 1], np.int32)
    with pytest.raises(SableError) as e:
        sp.SptsAnalysis(sp.csr_from_host(2, rp, c, np.array([1.0, 2.0, 0.0])))
    assert e.value.name == "SABLE_ESINGULAR"


def test_forced_spin_limit_reports_timeout(ctx):
    # a 4096-long dependency chain: every row waits on the previous one, so a
    # spin limit of one poll without progress must trip; the kernel exits.
    N = 4096
    rp = np.concatenate([[0, 1], 1 + 2 * np.arange(1, N)]).astype(np.int32)
    c = np.empty(2 * N - 1, np.int32)
    c[0] = 0
    c[1::2] = np.arange(N - 1)
    c[2::2] = np.arange(1, N)
    v = np.ones(2 * N - 1)
    L = sp.csr_from_host(N, rp, c, v)
    b = sp.create_vec(N, 1.0, np.float64)
    with pytest.raises(SableError) as e:
        sp.spts_syncfree(L, b, sp.create_vec(N, 0.0, np.float64), lanes=1, spin_limit=1)
    assert e.value.name == "SABLE_ETIMEOUT"
    x = sp.create_vec(N, 0.0, np.float64)
    sp.spts_syncfree(L, b, x, lanes=1)          # the default bound solves it
    assert np.array_equal(x.array(), oracle.spts_syncfree(N, rp, c, v, np.ones(N)))


def csc_of_lower(N, rp, col, val):
    rows = np.repeat(np.arange(N), np.diff(rp))
    order = np.lexsort((rows, col))
    cp = np.concatenate([[0], np.cumsum(np.bincount(col, minlength=N))]).astype(np.int32)
    return cp, rows[order].astype(np.int32), val[order]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("name", ["random_lower", "lap3d"])
def test_spts_csc_bit_exact(ctx, dtype, name):
    """The serial tree's column sweep (spts_csc_wrapper with inner_max = 2)
    on the GPU: bit-identical to the oracle's column sweep."""
    if name == "lap3d":
        rp, col, val = gen.laplacian_3d_lower(12, dtype)
    else:
        rp, col, val = gen.random_lower(5000, kmax=16, seed=71, dtype=dtype)
    N = len(rp) - 1
    cp, row, cval = csc_of_lower(N, rp, col, val)
    b = np.random.default_rng(72).uniform(-1, 1, N).astype(dtype)
    A = sp.csc_from_host(N, cp, row, cval)
    yv = sp.create_vec(N, 0.0, dtype)
    sp.spts_csc(A, sp.vec_from_host(b), yv, inner_max=2)
    yref = oracle.spts_csc(N, cp, row, cval, b)
    assert np.array_equal(yv.array().view(np.uint8), yref.view(np.uint8))
    # the analysis of the CSC input serves the other solves too
    an = sp.SptsAnalysis(A, lanes=1)
    x2 = sp.create_vec(N, 0.0, dtype)
    an.solve_syncfree(sp.vec_from_host(b), x2)
    assert np.array_equal(x2.array(), yref)


def test_spts_csc_rejects_diagonal_not_first(ctx):
    # column 0 holds rows (1, 0): its first entry is not the diagonal
    cp = np.array([0, 2, 3], np.int32)
    row = np.array([1, 0, 1], np.int32)
    A = sp.csc_from_host(2, cp, row, np.array([0.5, 2.0, 4.0]))
    with pytest.raises(SableError) as e:
        sp.spts_csc(A, sp.vec_from_host(np.ones(2)), sp.create_vec(2, 0.0, np.float64))
    assert e.value.name == "SABLE_EMALFORMED"


BLOCK_MATS = dict(MATS, lap3d_32=lambda: gen.laplacian_3d_lower(32),
                  random_lower_200k=lambda: gen.random_lower(200000, 32, seed=8),
                  random_lower_w8=lambda: gen.random_lower(30000, 8, seed=9),
                  single_row=lambda: (np.array([0, 1], np.int32), np.array([0], np.int32), np.array([2.0])))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("name", list(BLOCK_MATS))
def test_row_blocks_bit_exact(ctx, dtype, name):
    """The row-block schedule (one launch: a block's rows step through the
    levels with its solution in LDS, published values across blocks) keeps
    one thread per row and the stored entry order: its level solve equals
    the reference level solve and its sync-free solve the reference
    sync-free solve, bit for bit, for every block size, for schedules whose
    widest row takes 4, 8 and 16 entries ahead (and more than 16, in a
    second pass), and on a second solve with the same analysis."""
    rp, c, v = BLOCK_MATS[name]()
    v = v.astype(dtype)
    N = len(rp) - 1
    b = np.random.default_rng(11).uniform(-1, 1, N).astype(dtype)
    an = sp.SptsAnalysis(sp.csr_from_host(N, rp, c, v), lanes=4)
    assert an.schedule_info()["schedule"] == "blocks"
    want_l = host_level_solution(N, rp, c, v, b)
    want_s = oracle.spts_syncfree(N, rp, c, v, b)
    bv = sp.vec_from_host(b)
    for B in (64, 1024, 16384, 0):
        an.set_schedule("blocks", B)
        info = an.schedule_info()
        print(name, dtype.__name__, info, flush=True)
        for _ in range(2):
            xl = sp.create_vec(N, 0.0, dtype)
            an.solve_level(bv, xl)
            xs = sp.create_vec(N, 0.0, dtype)
            an.solve_syncfree(bv, xs)
            an.status()
            assert np.array_equal(xl.array().view(np.uint8), want_l.view(np.uint8)), (name, B)
            assert np.array_equal(xs.array().view(np.uint8), want_s.view(np.uint8)), (name, B)


def test_row_blocks_schedule_shape(ctx):
    """The 3D Laplacian in blocks of whole planes: a 16^3 grid in 256-row
    blocks has 16 blocks, each holding 31 levels (i + j + k for its k) of at
    most 16 rows: one step per level."""
    rp, c, v = gen.laplacian_3d_lower(16)
    an = sp.SptsAnalysis(sp.csr_from_host(len(rp) - 1, rp, c, v))
    an.set_schedule("blocks", 256)
    info = an.schedule_info()
    assert info["nblocks"] == 16 and info["nsteps"] == 16 * 31
    # every step holds its rows' off-diagonals, padded to the step's widest row
    assert info["slots"] >= len(c) - (len(rp) - 1)
    with pytest.raises(SableError):
        an.set_schedule("blocks", 3000)


@pytest.mark.parametrize("name", ["lap3d_32", "random_lower_20k"])
def test_tune_keeps_bits(ctx, name):
    """The tuner times the row-block sizes and the level graph at several
    lane counts and keeps the fastest; whichever it keeps, the level solve
    stays bit for bit the reference level solve."""
    rp, c, v = BLOCK_MATS[name]()
    N = len(rp) - 1
    b = np.random.default_rng(12).uniform(-1, 1, N)
    an = sp.SptsAnalysis(sp.csr_from_host(N, rp, c, v))
    info = an.tune()
    print(name, info, flush=True)
    assert info["schedule"] in ("blocks", "graph") and info["ms"] > 0
    assert 1024 <= info["block_rows"] <= 16384 and info["lanes"] in (1, 4, 8, 16, 64)
    x = sp.create_vec(N, 0.0, np.float64)
    an.solve_level(sp.vec_from_host(b), x)
    an.status()
    assert np.array_equal(x.array(), host_level_solution(N, rp, c, v, b))


@pytest.mark.parametrize("lanes", [1, 4, 16])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("name", list(BLOCK_MATS))
def test_indegree_syncfree_bit_exact(ctx, lanes, dtype, name):
    """The sync-free solve with per-row in-degree counters: the reference
    sync-free solve's bits at every lane count, on two solves in a row (the
    counters are reset by each launch)."""
    rp, c, v = BLOCK_MATS[name]()
    v = v.astype(dtype)
    N = len(rp) - 1
    b = np.random.default_rng(13).uniform(-1, 1, N).astype(dtype)
    want = oracle.spts_syncfree(N, rp, c, v, b)
    an = sp.SptsAnalysis(sp.csr_from_host(N, rp, c, v), lanes=lanes)
    bv = sp.vec_from_host(b)
    for _ in range(2):
        x = sp.create_vec(N, 0.0, dtype)
        an.solve_indegree(bv, x)
        an.status()
        assert np.array_equal(x.array().view(np.uint8), want.view(np.uint8)), name


def test_graph_lane_changes_without_wait(ctx):
    """Lanes changed between queued graph solves, with no host wait between
    them: each change re-captures the level graph, and the old graph's
    queued replay still finishes first; every solve gives the reference
    level solve's bits."""
    rp, c, v = gen.random_lower(40000, 16, seed=11)
    N = len(rp) - 1
    b = np.random.default_rng(12).uniform(-1, 1, N)
    an = sp.SptsAnalysis(sp.csr_from_host(N, rp, c, v), lanes=1)
    an.set_schedule("graph")
    bv = sp.vec_from_host(b)
    seq = [1, 4, 8, 16, 4, 1]
    xs = [sp.create_vec(N, 0.0, np.float64) for _ in seq]
    ctx.sync()
    for lanes, x in zip(seq, xs):
        an.set_lanes(lanes)
        an.solve_level(bv, x)
    want = host_level_solution(N, rp, c, v, b)
    for lanes, x in zip(seq, xs):
        assert np.array_equal(x.array(), want), lanes


def host_fused_refine(N, rp, c, v, b):
    """The fused refined solve computed exactly on the host: x0 the FP32
    reference level solve of (float)b; r = (float)(b - s) with s the FP64
    products of the row chained by fused multiply-add in stored order
    (diagonal last; each fma rounded once, computed with exact rationals);
    d the FP32 level solve of r; x = x0 + d in FP64."""
    from fractions import Fraction
    v32 = v.astype(np.float32)
    x0 = host_level_solution(N, rp, c, v32, b.astype(np.float32))
    r = np.empty(N, np.float32)
    for i in range(N):
        s = 0.0
        for j in range(rp[i], rp[i + 1]):
            s = float(Fraction(float(v[j])) * Fraction(float(x0[c[j]])) + Fraction(s))
        r[i] = np.float32(b[i] - s)
    d = host_level_solution(N, rp, c, v32, r)
    return x0.astype(np.float64) + d.astype(np.float64)


@pytest.mark.parametrize("name", ["lap3d_16", "random_lower_sparse", "random_lower_20k"])
def test_fused_refinement_every_schedule(ctx, name):
    """FP32 solve + one FP64 refinement step in one sweep: the same x on the
    row blocks (two block sizes), the level graph and the per-level launches
    at 1, 4 and 16 lanes, equal to the exact host computation of the same
    steps, within 1e-12 of the FP64 oracle, and no slower to converge than
    the unfused refinement."""
    rp, c, v = MATS[name]()
    N = len(rp) - 1
    b = np.random.default_rng(21).uniform(-1, 1, N)
    an = sp.SptsAnalysis(sp.csr_from_host(N, rp, c, v))
    bv = sp.vec_from_host(b)
    want = host_fused_refine(N, rp, c, v, b)
    runs = []
    for sched, arg in (("blocks", 1024), ("blocks", 4096), ("graph", 1), ("graph", 4), ("graph", 16),
                       ("launches", 4)):
        if sched == "blocks":
            an.set_schedule("blocks", arg)
        else:
            an.set_lanes(arg)
            an.set_schedule(sched)
        for _ in range(2):
            x = sp.create_vec(N, 0.0, np.float64)
            an.solve_refined(bv, x, iters=1, inner="fused")
            an.status()
            runs.append(((sched, arg), x.array()))
    for key, xa in runs:
        assert np.array_equal(xa, want), key
    assert oracle.rel_residual(N, rp, c, v, want, b) <= 1e-12
    x = sp.create_vec(N, 0.0, np.float64)
    an.solve_refined(bv, x, iters=1, inner="level")
    an.status()
    assert oracle.rel_residual(N, rp, c, v, x.array(), b) <= 1e-12


def test_fused_refinement_needs_one_step(ctx):
    rp, c, v = MATS["lap3d_16"]()
    N = len(rp) - 1
    an = sp.SptsAnalysis(sp.csr_from_host(N, rp, c, v))
    bv = sp.vec_from_host(np.ones(N))
    for iters in (0, 2):
        with pytest.raises(SableError) as e:
            an.solve_refined(bv, sp.create_vec(N, 0.0, np.float64), iters=iters, inner="fused")
        assert e.value.name == "SABLE_EINVAL"
