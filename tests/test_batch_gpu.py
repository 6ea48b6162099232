"""Batched persistent launches against per-matrix launches and the oracle.

* batch_spmv with one right-hand side runs the single-matrix block kernel's
  arithmetic on the same row-block plan, so it must equal the per-matrix
  launch bit for bit; with nrhs > 1 each right-hand side must meet the CSR
  SpMV bound of test_spmv_gpu.py against the oracle.
* batch_spts at every lane count is bit-identical to the reference-order
  sync-free solve (lanes form a row's products in parallel and subtract them
  in stored order); the FP32 + FP64-refinement variant must reach relative
  residual <= 1e-12 against the long-double oracle."""
import numpy as np
import pytest

import oracle
import sciwasm.sparse_amd as sp
from sciwasm.sparse_amd import gen, DeviceArray, SableError

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mats():
    return gen.batch_csr(40, rows_lo=512, rows_hi=6000, seed=11)


def bound(rp, col, val, x, dtype):
    u = np.finfo(dtype).eps / 2
    k = np.diff(rp).astype(np.float64) + 1
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    s = np.bincount(rows, weights=np.abs(val.astype(np.float64) * x.astype(np.float64)[col]), minlength=len(rp) - 1)
    return 2 * k * u / (1 - k * u) * s + np.finfo(dtype).tiny


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_batch_spmv_equals_single(ctx, mats, dtype):
    As, xs, ys, ref = [], [], [], []
    rng = np.random.default_rng(1)
    for n, rp, c, v in mats:
        A = sp.csr_from_host(n, rp, c, v.astype(dtype))
        x = rng.uniform(-1, 1, n).astype(dtype)
        As.append(A)
        xs.append(DeviceArray.from_host(x))
        ys.append(DeviceArray.zeros(n, dtype))
        yv = sp.create_vec(n, 0.0, dtype)
        sp.spmv(A, sp.vec_from_host(x), yv, beta=0.0)
        ref.append(yv.array())
    sp.batch_spmv(As, xs, ys, nrhs=1)
    for k in range(len(As)):
        assert np.array_equal(ys[k].numpy(), ref[k]), k


@pytest.mark.parametrize("nrhs", [2, 4, 8])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_batch_spmv_multi_rhs(ctx, mats, nrhs, dtype):
    rng = np.random.default_rng(2)
    As, xs, ys, X = [], [], [], []
    for n, rp, c, v in mats[:12]:
        As.append(sp.csr_from_host(n, rp, c, v.astype(dtype)))
        x = rng.uniform(-1, 1, (n, nrhs)).astype(dtype)
        X.append(x)
        xs.append(DeviceArray.from_host(x.ravel()))
        ys.append(DeviceArray.zeros(n * nrhs, dtype))
    sp.batch_spmv(As, xs, ys, nrhs=nrhs)
    for k, (n, rp, c, v) in enumerate(mats[:12]):
        Y = ys[k].numpy().reshape(n, nrhs)
        for r in range(nrhs):
            yref = oracle.spmv_csr(n, rp, c, v.astype(dtype), X[k][:, r].copy(), np.zeros(n, dtype))
            d = np.abs(Y[:, r].astype(np.float64) - yref)
            assert np.all(d <= bound(rp, c, v.astype(dtype), X[k][:, r], dtype)), (k, r)


@pytest.mark.parametrize("nrhs", [1, 8])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_batch_spmv_blocks_beyond_lds(ctx, mats, nrhs, dtype):
    """Plans with blocks of up to 4096 weight: blocks whose products do not
    fit the LDS staging buffer take the wave-per-row path, one-row blocks
    the whole-workgroup path; results within the CSR bound."""
    rng = np.random.default_rng(3)
    As, plans, xs, ys, X = [], [], [], [], []
    sub = mats[:10]
    for n, rp, c, v in sub:
        A = sp.csr_from_host(n, rp, c, v.astype(dtype))
        As.append(A)
        plans.append(sp.CsrPlan(A, {"block_weight": 4096, "row_weight": 1}))
        x = rng.uniform(-1, 1, (n, nrhs)).astype(dtype)
        X.append(x)
        xs.append(DeviceArray.from_host(x.ravel()))
        ys.append(DeviceArray.zeros(n * nrhs, dtype))
    sp.batch_spmv(As, xs, ys, nrhs=nrhs, plans=plans)
    for k, (n, rp, c, v) in enumerate(sub):
        Y = ys[k].numpy().reshape(n, nrhs)
        for r in range(nrhs):
            yref = oracle.spmv_csr(n, rp, c, v.astype(dtype), X[k][:, r].copy(), np.zeros(n, dtype))
            d = np.abs(Y[:, r].astype(np.float64) - yref)
            assert np.all(d <= bound(rp, c, v.astype(dtype), X[k][:, r], dtype)), (k, r)


def tiles_case(dtype):
    """Uneven matrices for the batched tile launch: the batch law plus sizes
    at and around the 64-row and 4096-row tile edges, a matrix with empty
    rows and an empty matrix."""
    out = [(n, rp, c, v.astype(dtype)) for n, rp, c, v in gen.batch_csr(16, rows_lo=512, rows_hi=20000, seed=21)]
    for k, n in enumerate((1, 63, 64, 65, 4095, 4097, 9000)):
        rp, c, v = gen.random_csr(n, 9, seed=40 + k, dtype=dtype)
        out.append((n, rp, c, v))
    rp, c, v = gen.random_csr(7000, 6, seed=50, dtype=dtype, empty_rows=0.3)
    out.append((7000, rp, c, v))
    out.append((0, np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0, dtype)))
    return out


@pytest.mark.parametrize("beta", [0.0, 1.0, 0.5])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_batch_spmv_tiles(ctx, dtype, beta):
    """batch_spmv_tiles against the oracle: every matrix's rows within the
    CSR bound (the LDS sums may add in any order), plus one rounding of
    beta*y0 + sum for beta != 0."""
    rng = np.random.default_rng(5)
    case = tiles_case(dtype)
    As, xs, ys, X, Y0 = [], [], [], [], []
    for n, rp, c, v in case:
        A = sp.csr_from_host(n, rp, c, v) if n else None
        As.append(A)
        x = rng.uniform(-1, 1, n).astype(dtype)
        y0 = rng.uniform(-1, 1, n).astype(dtype)
        X.append(x)
        Y0.append(y0)
        xs.append(DeviceArray.from_host(x) if n else None)
        ys.append(DeviceArray.from_host(y0) if n else None)
    made = iter(sp.batch_tile_plans([A for A in As if A is not None]))
    plans = [next(made) if A is not None else None for A in As]
    sp.batch_spmv_tiles(plans, xs, ys, beta=beta)
    ctx.sync()
    u = np.finfo(dtype).eps / 2
    for k, (n, rp, c, v) in enumerate(case):
        if n == 0:
            assert plans[k] is None
            continue
        yref = oracle.spmv_csr(n, rp, c, v, X[k].copy(), np.zeros(n, dtype)).astype(np.float64)
        b = bound(rp, c, v, X[k], dtype)
        by0 = beta * Y0[k].astype(np.float64)
        want = by0 + yref
        tol = b + 2 * u * (np.abs(by0) + np.abs(yref) + b)
        d = np.abs(ys[k].numpy().astype(np.float64) - want)
        assert np.all(d <= tol), (k, n, float(np.max(d - tol)))


@pytest.mark.parametrize("few_units", [False, True], ids=["all", "under8units"])
@pytest.mark.parametrize("workgroups", [1, 3, 7])
def test_batch_spmv_tiles_small_grids(ctx, workgroups, few_units):
    """Grids of fewer workgroups than claim ranges (8): the ranges without a
    home workgroup are drained by stealing, and a batch of fewer than 8
    units is complete; every matrix within the CSR bound."""
    dtype = np.float64
    rng = np.random.default_rng(6)
    case = [c for c in tiles_case(dtype) if c[0] > 0]
    if few_units:
        case = [c for c in case if c[0] <= 4096][:3]  # one row tile each
    As = [sp.csr_from_host(n, rp, c, v) for n, rp, c, v in case]
    X = [rng.uniform(-1, 1, n).astype(dtype) for n, _, _, _ in case]
    xs = [DeviceArray.from_host(x) for x in X]
    ys = [DeviceArray.zeros(n, dtype) for n, _, _, _ in case]
    plans = sp.batch_tile_plans(As)
    total = sum(p.info()["row_tiles"] for p in plans)
    if few_units:
        assert total < 8
    sp.batch_spmv_tiles(plans, xs, ys, workgroups=workgroups)
    ctx.sync()
    for k, (n, rp, c, v) in enumerate(case):
        yref = oracle.spmv_csr(n, rp, c, v, X[k], np.zeros(n, dtype))
        d = np.abs(ys[k].numpy() - yref)
        assert np.all(d <= bound(rp, c, v, X[k], dtype)), (workgroups, k)


@pytest.mark.parametrize("nrhs,beta", [(2, 0.0), (4, 0.5), (8, 0.0), (8, 1.0)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_batch_spmv_tiles_multi_rhs(ctx, dtype, nrhs, beta):
    """nrhs right-hand sides, x and y row-major [N][nrhs]: every column
    within the CSR bound of its own product (plus one rounding of
    beta*y0 + sum), with the default tiles and, for 4, the largest that
    the LDS holds (4096 rows, 128 KiB of row sums)."""
    rng = np.random.default_rng(7 + nrhs)
    case = [c for c in tiles_case(dtype) if c[0] > 0]
    As = [sp.csr_from_host(n, rp, c, v) for n, rp, c, v in case]
    plans = sp.batch_tile_plans(As, tile_rows=4096 if nrhs == 4 else 0, nrhs=nrhs)
    X = [rng.uniform(-1, 1, (n, nrhs)).astype(dtype) for n, _, _, _ in case]
    Y0 = [rng.uniform(-1, 1, (n, nrhs)).astype(dtype) for n, _, _, _ in case]
    xs = [DeviceArray.from_host(x.ravel()) for x in X]
    ys = [DeviceArray.from_host(y.ravel()) for y in Y0]
    sp.batch_spmv_tiles(plans, xs, ys, nrhs=nrhs, beta=beta)
    ctx.sync()
    u = np.finfo(dtype).eps / 2
    for k, (n, rp, c, v) in enumerate(case):
        Y = ys[k].numpy().reshape(n, nrhs).astype(np.float64)
        for r in range(nrhs):
            x = X[k][:, r].copy()
            yref = oracle.spmv_csr(n, rp, c, v, x, np.zeros(n, dtype)).astype(np.float64)
            b = bound(rp, c, v, x, dtype)
            by0 = beta * Y0[k][:, r].astype(np.float64)
            tol = b + 2 * u * (np.abs(by0) + np.abs(yref) + b)
            d = np.abs(Y[:, r] - (by0 + yref))
            assert np.all(d <= tol), (k, n, r)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_batch_spmv_tiles_many_panels(ctx, dtype):
    """Plans with small tiles and several panels (the general tile layout)
    in one batch, equal to each plan's own one-launch product within the
    bound, and max_tile_rows below an item's tile_rows refused after a
    launch that skips that item."""
    import ctypes
    from sciwasm.sparse_amd import _lib
    from sciwasm.sparse_amd.serve import BatchTilesItem, _upload_items
    from sciwasm.sparse_amd.core import dtype_code
    rng = np.random.default_rng(6)
    shapes = [(3000, 128, 256), (5000, 256, 1024), (777, 64, 64), (12000, 1024, 2048)]
    As, plans, xs, ys, X, M = [], [], [], [], [], []
    for k, (n, tr, pc) in enumerate(shapes):
        rp, c, v = gen.random_csr(n, 11, seed=60 + k, dtype=dtype)
        M.append((n, rp, c, v))
        A = sp.csr_from_host(n, rp, c, v)
        As.append(A)
        plans.append(sp.CsrTiles(A, tile_rows=tr, panel_cols=pc))
        x = rng.uniform(-1, 1, n).astype(dtype)
        X.append(x)
        xs.append(DeviceArray.from_host(x))
        ys.append(DeviceArray.zeros(n, dtype))
    sp.batch_spmv_tiles(plans, xs, ys)
    ctx.sync()
    for k, (n, rp, c, v) in enumerate(M):
        yref = oracle.spmv_csr(n, rp, c, v, X[k].copy(), np.zeros(n, dtype))
        d = np.abs(ys[k].numpy().astype(np.float64) - yref)
        assert np.all(d <= bound(rp, c, v, X[k], dtype)), k
    items, total = [], 0
    for p, x, y in zip(plans, xs, ys):
        it = BatchTilesItem()
        _lib.call("sable_csr_tiles_batch_item", p.handle, x.ptr, y.ptr, 0.0, total, ctypes.byref(it))
        items.append(it)
        total += it.ntr
    d = _upload_items(items, ctx)
    with pytest.raises(SableError) as e:
        _lib.call("sable_batch_spmv_tiles", ctx.handle, dtype_code(dtype), 1, d.ptr, len(items), total, 512, 0)
    assert e.value.name == "SABLE_EINVAL"
    # 8 right-hand sides of 4096-row tiles need 256 KiB of row sums: refused before any launch
    with pytest.raises(SableError) as e:
        _lib.call("sable_batch_spmv_tiles", ctx.handle, dtype_code(dtype), 8, d.ptr, len(items), total, 4096, 0)
    assert e.value.name == "SABLE_EINVAL"


@pytest.mark.parametrize("allow_tiles", [True, False], ids=["tiles", "rowblock_only"])
@pytest.mark.parametrize("nrhs", [1, 4])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_autotuned_batch(ctx, dtype, nrhs, allow_tiles):
    """The autotuned batch (row-length classes, each with the fastest of
    row-block plans of several block/row weights and tile plans of several
    tile heights, measured on the class's own matrices in batched launches;
    then the fastest whole-batch assignment of that and the uniform ones)
    over matrices of 64 to 20 000 rows: every product within the CSR bound
    of the oracle, twice in a row (the second with beta = 1 on the first's
    y), and without tiles the row-block batch's bits (each row in a fixed
    order) on a repeat."""
    rng = np.random.default_rng(9)
    mm = gen.batch_csr(30, rows_lo=64, rows_hi=20000, seed=12) + [(0, np.zeros(1, np.int32),
                                                                 np.zeros(0, np.int32), np.zeros(0))]
    As, xs, ys, X = [], [], [], []
    for n, rp, c, v in mm:
        As.append(sp.csr_from_host(n, rp, c, v.astype(dtype)))
        x = rng.uniform(-1, 1, (n, nrhs)).astype(dtype)
        X.append(x)
        xs.append(DeviceArray.from_host(x.ravel()) if n else DeviceArray.zeros(1, dtype))
        ys.append(DeviceArray.zeros(max(n, 1) * nrhs, dtype))
    bp = sp.BatchPlan(As, nrhs=nrhs, reps=2, allow_tiles=allow_tiles)
    info = bp.info()
    print(dtype.__name__, nrhs, allow_tiles, {k: (v["format"], v["matrices"]) for k, v in info.items()}, flush=True)
    assert sum(v["matrices"] for v in info.values()) == len(mm) - 1
    assert allow_tiles or all(v["format"] == "rowblock" for v in info.values())
    ib = bp.info_batch()
    names = {t["assignment"] for t in ib["table"]}
    assert ib["assignment"] in names and "per_class" in names and "all_rowblock" in names
    assert ib["ms"] == min(t["ms"] for t in ib["table"])
    assert allow_tiles or names == {"per_class", "all_rowblock"}
    bp.run(xs, ys)
    ctx.sync()
    first = [y.numpy().copy() for y in ys]
    for k, (n, rp, c, v) in enumerate(mm):
        if n == 0:
            continue
        Y = first[k][:n * nrhs].reshape(n, nrhs)
        for r in range(nrhs):
            yref = oracle.spmv_csr(n, rp, c, v.astype(dtype), X[k][:, r].copy(), np.zeros(n, dtype))
            d = np.abs(Y[:, r].astype(np.float64) - yref)
            assert np.all(d <= bound(rp, c, v.astype(dtype), X[k][:, r], dtype)), (k, r)
    run = bp.bind(xs, ys, beta=1.0)
    run.run()
    ctx.sync()
    u = np.finfo(dtype).eps / 2
    for k, (n, rp, c, v) in enumerate(mm):
        if n == 0:
            continue
        Y = ys[k].numpy()[:n * nrhs].reshape(n, nrhs)
        Y0 = first[k][:n * nrhs].reshape(n, nrhs).astype(np.float64)
        for r in range(nrhs):
            yref = oracle.spmv_csr(n, rp, c, v.astype(dtype), X[k][:, r].copy(), np.zeros(n, dtype))
            b = bound(rp, c, v.astype(dtype), X[k][:, r], dtype)
            assert np.all(np.abs(Y[:, r] - (Y0[:, r] + yref)) <= 2 * b + 2 * u * (np.abs(Y0[:, r]) + np.abs(yref)))
    if not allow_tiles:
        for y in ys:
            y.fill(0.0)
        bp.run(xs, ys)
        ctx.sync()
        for k, (n, rp, c, v) in enumerate(mm):
            assert np.array_equal(ys[k].numpy(), first[k]), k


def lower_set(mats, dtype=np.float64):
    out = []
    for n, rp, c, v in mats:
        lrp, lc, lv = sp.lower_dominant(rp, c, v)
        out.append((n, lrp, lc, lv.astype(dtype)))
    return out


@pytest.mark.parametrize("lanes", [1, 4, 16])
def test_batch_spts(ctx, mats, lanes):
    Ls = lower_set(mats)
    rng = np.random.default_rng(3)
    dL, bs, xs, bh = [], [], [], []
    for n, rp, c, v in Ls:
        b = rng.uniform(-1, 1, n)
        bh.append(b)
        dL.append(sp.csr_from_host(n, rp, c, v))
        bs.append(DeviceArray.from_host(b))
        xs.append(DeviceArray.zeros(n, np.float64))
    sp.batch_spts(dL, bs, xs, lanes=lanes)
    for k, (n, rp, c, v) in enumerate(Ls):
        x = xs[k].numpy()
        assert np.array_equal(x, oracle.spts_syncfree(n, rp, c, v, bh[k])), k


def level_solution(n, rp, c, v, b):
    nl, lp, perm, order, _ = oracle.level_sets(n, rp, c)
    nrp, nc_, nv = oracle.level_reorder(n, rp, c, v, perm, order)
    yl = oracle.spts_level(n, nrp, nc_, nv, b[perm])
    x = np.empty_like(yl)
    x[perm] = yl
    return x


@pytest.mark.parametrize("lanes", [1, 4, 16])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_batch_level_plan(ctx, mats, lanes, dtype):
    """The batched level schedule: every item's solve is the reference level
    solve's bits, twice with other right-hand sides; one matrix with only
    its diagonal in the batch."""
    Ls = lower_set(mats, dtype) + [(3, np.arange(4, dtype=np.int32), np.arange(3, dtype=np.int32),
                                    np.array([2.0, 4.0, 8.0], dtype))]
    dL = [sp.csr_from_host(n, rp, c, v) for n, rp, c, v in Ls]
    plan = sp.BatchLevelPlan(dL, lanes=lanes)
    info = plan.info()
    assert info["levels"] >= 1 and info["chunks"] >= len(Ls)
    rng = np.random.default_rng(21)
    for rep in range(2):
        bh = [rng.uniform(-1, 1, n).astype(dtype) for n, _, _, _ in Ls]
        bs = [DeviceArray.from_host(b) for b in bh]
        xs = [DeviceArray.zeros(n, dtype) for n, _, _, _ in Ls]
        plan.solve(bs, xs)
        ctx.sync()
        for k, (n, rp, c, v) in enumerate(Ls):
            assert np.array_equal(xs[k].numpy(), level_solution(n, rp, c, v, bh[k])), (rep, k)


@pytest.mark.parametrize("lanes", [1, 4, 16])
def test_batch_level_plan_refined(ctx, mats, lanes):
    """The batched level schedule with the FP32 solve and one FP64
    refinement step fused into its sweep: each item's x is the exact host
    computation of those steps (the single-matrix fused solve's x; checked on
    a sample of items) and every item's is within 1e-12 of the FP64 oracle,
    twice with other right-hand sides."""
    from test_spts_gpu import host_fused_refine
    Ls = lower_set(mats) + [(3, np.arange(4, dtype=np.int32), np.arange(3, dtype=np.int32),
                             np.array([2.0, 4.0, 8.0]))]
    dL = [sp.csr_from_host(n, rp, c, v) for n, rp, c, v in Ls]
    plan = sp.BatchLevelPlan(dL, lanes=lanes)
    rng = np.random.default_rng(23)
    for rep in range(2):
        bh = [rng.uniform(-1, 1, n) for n, _, _, _ in Ls]
        bs = [DeviceArray.from_host(b) for b in bh]
        xs = [DeviceArray.zeros(n, np.float64) for n, _, _, _ in Ls]
        plan.bind_refined(bs, xs).run()
        ctx.sync()
        for k, (n, rp, c, v) in enumerate(Ls):
            x = xs[k].numpy()
            if k < 6 or k == len(Ls) - 1:  # the exact host steps are slow: a sample
                assert np.array_equal(x, host_fused_refine(n, rp, c, v, bh[k])), (rep, k)
            assert oracle.rel_residual(n, rp, c, v, x, bh[k]) <= 1e-12, (rep, k)


def test_batch_level_plan_rejects_malformed(ctx):
    rp = np.array([0, 1, 3], np.int32)
    c = np.array([0, 1, 0], np.int32)  # row 1 ends with column 0: the diagonal is not last
    A = sp.csr_from_host(2, rp, c, np.ones(3))
    with pytest.raises(SableError) as e:
        sp.BatchLevelPlan([A])
    assert e.value.name == "SABLE_EMALFORMED"


def test_batch_spts_refined(ctx, mats):
    Ls = lower_set(mats)
    rng = np.random.default_rng(4)
    dL, bs, xs, bh = [], [], [], []
    for n, rp, c, v in Ls:
        b = rng.uniform(-1, 1, n)
        bh.append(b)
        dL.append(sp.csr_from_host(n, rp, c, v))
        bs.append(DeviceArray.from_host(b))
        xs.append(DeviceArray.zeros(n, np.float64))
    sp.batch_spts_refined(dL, bs, xs, lanes=4, iters=2)
    for k, (n, rp, c, v) in enumerate(Ls):
        assert oracle.rel_residual(n, rp, c, v, xs[k].numpy(), bh[k]) <= 1e-12, k
