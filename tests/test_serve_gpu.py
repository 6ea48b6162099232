"""The request server: results alone and in a concurrent mix, accuracy
against the oracle, HBM-budget admission, and a short open-loop run.

Alone vs in the mix: every mode and solve sweep computes each request with
an arithmetic that does not depend on its batch (the batched kernels keep each item's
per-row order; the solves' per-row order is fixed by the lane layout), so
the two results must be identical; the stated tolerance against the oracle
is the CSR SpMV bound for SpMV (FP32 unit roundoff in mixed mode) and
relative residual <= 1e-12 for the solves in every mode."""
import numpy as np
import pytest

import oracle
import sciwasm.sparse_amd as sp
from sciwasm.sparse_amd import gen, SableError

from test_spts_gpu import host_fused_refine

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pool():
    mats = gen.batch_csr(24, rows_lo=512, rows_hi=8000, seed=21)
    lows = [(n,) + sp.lower_dominant(rp, c, v) for n, rp, c, v in mats]
    return mats, lows


def make_server(mode, pool, **kw):
    mats, lows = pool
    S = sp.Server(mode=mode, arena_bytes=64 << 20, **kw)
    gids = [S.add_matrix(rp, c, v) for n, rp, c, v in mats]
    lids = [S.add_matrix(rp, c, v, lower=True) for n, rp, c, v in lows]
    return S, gids, lids


def spmv_bound(rp, col, val, x, dtype):
    u = np.finfo(dtype).eps / 2
    k = np.diff(rp).astype(np.float64) + 1
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    s = np.bincount(rows, weights=np.abs(val * x[col]), minlength=len(rp) - 1)
    # mixed mode also rounds x and the values to FP32: 2u more per product
    return (2 * k + 4) * u / (1 - (k + 2) * u) * s + 1e-300


def level_solution(n, rp, c, v, b):
    nl, lp, perm, order, _ = oracle.level_sets(n, rp, c)
    nrp, nc_, nv = oracle.level_reorder(n, rp, c, v, perm, order)
    yl = oracle.spts_level(n, nrp, nc_, nv, b[perm])
    x = np.empty_like(yl)
    x[perm] = yl
    return x


@pytest.mark.parametrize("mode,kw", [("plain", {}), ("batched", {}), ("batched", {"group_rhs": 1}), ("mixed", {}),
                                     ("mixed", {"refine_iters": 2}), ("batched", {"spts_sweep": "syncfree"}),
                                     ("mixed", {"spts_sweep": "syncfree"})],
                         ids=["plain", "batched", "batched_ungrouped", "mixed", "mixed_refine2", "batched_sf",
                              "mixed_sf"])
def test_alone_equals_mix_and_accuracy(ctx, pool, mode, kw):
    """Also: the FP64 solves are a reference solve's bits, the sync-free
    solve's (plain; batched one-launch sweep) or the level solve's (batched
    level sweep); the mixed solves with one refinement step (the default,
    fused into the level sweep) are the exact host computation of those
    steps."""
    mats, lows = pool
    S, gids, lids = make_server(mode, pool, **kw)
    rng = np.random.default_rng(5)
    probes = []
    for k in range(6):
        n, rp, c, v = mats[k]
        probes.append(("spmv", gids[k], rng.uniform(-1, 1, n), (n, rp, c, v)))
        n, rp, c, v = lows[k]
        probes.append(("spts", lids[k], rng.uniform(-1, 1, n), (n, rp, c, v)))
    alone = []
    for op, mid, vec, _ in probes:
        out, lat = S.wait(S.submit(mid, op, vec))
        alone.append(out.copy())
    # the same requests interleaved with 200 others submitted at once (SpMV
    # requests on one matrix then share multi-RHS products)
    ids = []
    for j in range(200):
        k = rng.integers(len(gids))
        if j % 3 == 0:
            ids.append(S.submit(lids[k], "spts", rng.uniform(-1, 1, lows[k][0])))
        else:
            ids.append(S.submit(gids[k], "spmv", rng.uniform(-1, 1, mats[k][0])))
    probe_ids = {}
    for p, (op, mid, vec, _) in enumerate(probes):
        probe_ids[p] = S.submit(mid, op, vec)
        for j in range(10):
            k = rng.integers(len(gids))
            ids.append(S.submit(gids[k], "spmv", rng.uniform(-1, 1, mats[k][0])))
    for p, (op, mid, vec, (n, rp, c, v)) in enumerate(probes):
        out, lat = S.wait(probe_ids[p])
        assert np.array_equal(out, alone[p]), (mode, p, op)
        if op == "spmv":
            yref, _ = oracle.spmv_csr_fp64(n, rp, c, v, vec)
            dt = np.float32 if mode == "mixed" else np.float64
            assert np.all(np.abs(out - yref) <= spmv_bound(rp, c, v, vec, dt)), (mode, p)
        else:
            assert oracle.rel_residual(n, rp, c, v, out, vec) <= 1e-12, (mode, p)
            if mode == "plain" or (mode == "batched" and kw.get("spts_sweep") == "syncfree"):
                assert np.array_equal(out, oracle.spts_syncfree(n, rp, c, v, vec)), (mode, p)
            elif mode == "batched":
                assert np.array_equal(out, level_solution(n, rp, c, v, vec)), (mode, p)
            elif not kw:
                assert np.array_equal(out, host_fused_refine(n, rp, c, v, vec)), (mode, p)
    for i in ids:
        S.wait(i)
    S.close()


@pytest.mark.parametrize("mode", ["batched", "mixed"])
def test_batch_tiles_alone_and_in_mix(ctx, pool, mode):
    """batch_tiles: the pool's SpMV runs as column-sorted tile plans in one
    batched launch per width, rows summed in LDS-atomic order.  Alone and in
    a mix (multi-RHS groups included) each SpMV result meets the CSR bound
    against the FP64 oracle, so the two agree within twice that bound; the
    solves are untouched and stay bit-identical."""
    mats, lows = pool
    S, gids, lids = make_server(mode, pool, batch_tiles=True)
    rng = np.random.default_rng(15)
    probes = []
    for k in range(6):
        n, rp, c, v = mats[k]
        probes.append(("spmv", gids[k], rng.uniform(-1, 1, n), (n, rp, c, v)))
        n, rp, c, v = lows[k]
        probes.append(("spts", lids[k], rng.uniform(-1, 1, n), (n, rp, c, v)))
    alone = [S.wait(S.submit(mid, op, vec))[0].copy() for op, mid, vec, _ in probes]
    ids = []
    for j in range(120):
        k = rng.integers(len(gids))
        ids.append(S.submit(gids[k], "spmv", rng.uniform(-1, 1, mats[k][0])))
    probe_ids = [S.submit(mid, op, vec) for op, mid, vec, _ in probes]
    for k in range(6):
        ids.append(S.submit(gids[k], "spmv", rng.uniform(-1, 1, mats[k][0])))
    dt = np.float32 if mode == "mixed" else np.float64
    for p, (op, mid, vec, (n, rp, c, v)) in enumerate(probes):
        out, _ = S.wait(probe_ids[p])
        if op == "spmv":
            yref, _ = oracle.spmv_csr_fp64(n, rp, c, v, vec)
            b = spmv_bound(rp, c, v, vec, dt)
            assert np.all(np.abs(out - yref) <= b), (mode, p)
            assert np.all(np.abs(alone[p] - yref) <= b), (mode, p)
            assert np.all(np.abs(out - alone[p]) <= 2 * b), (mode, p)
        else:
            assert np.array_equal(out, alone[p]), (mode, p)
            assert oracle.rel_residual(n, rp, c, v, out, vec) <= 1e-12, (mode, p)
    for i in ids:
        S.wait(i)
    S.close()


def test_hbm_budget_admission(ctx, pool):
    mats, lows = pool
    # per slot: the 16 MiB arena, 4 MiB of item arrays, 8 MiB of level
    # segments and 4 KiB of counters; 4 MiB left for matrices
    S = sp.Server(mode="batched", hbm_budget=3 * ((16 << 20) + (4 << 20) + (8 << 20) + 4096) + (4 << 20),
                  arena_bytes=16 << 20, slots=3)
    n, rp, c, v = mats[0]
    S.add_matrix(rp, c, v)
    big = gen.band_csr(200000, [-2, -1, 0, 1, 2], seed=1)
    with pytest.raises(SableError) as e:
        S.add_matrix(*big)
    assert e.value.name == "SABLE_EBUDGET"
    S.close()
    # a request that can never fit a batch arena is rejected, not hung
    S = sp.Server(mode="batched", arena_bytes=1 << 20)
    mid = S.add_matrix(*big)
    with pytest.raises(SableError) as e:
        S.wait(S.submit(mid, "spmv", np.ones(200000)))
    assert e.value.name == "SABLE_EBUDGET"
    S.close()


def test_malformed_lower_rejected(ctx):
    S = sp.Server(mode="batched", arena_bytes=1 << 20)
    with pytest.raises(SableError) as e:
        S.add_matrix(np.array([0, 2, 3], np.int32), np.array([1, 0, 1], np.int32), np.ones(3), lower=True)
    assert e.value.name == "SABLE_EMALFORMED"
    S.close()


@pytest.mark.parametrize("mode,sweep", [("plain", "levels"), ("mixed", "levels"), ("mixed", "syncfree")])
def test_open_loop_load(ctx, pool, mode, sweep):
    S, gids, lids = make_server(mode, pool, spts_sweep=sweep)
    st = S.run_load(rate=2000, duration_s=0.5, spts_fraction=0.3, seed=7)
    assert st["failed"] == 0 and st["rejected"] == 0
    assert st["completed"] == st["submitted"] > 0
    assert 0 < st["p50_us"] <= st["p99_us"] <= st["max_us"]
    if mode == "mixed":
        assert st["mean_batch"] >= 1.0
    S.close()


def test_grouped_requests_on_one_matrix(ctx, pool):
    """Nine SpMV requests on one matrix submitted together (they share
    multi-RHS products, of widths up to 8, with whichever requests the
    scheduler admits into the same batch); each result equals the request
    computed alone."""
    mats, _ = pool
    S = sp.Server(mode="batched", arena_bytes=64 << 20, max_batch=64)
    n, rp, c, v = mats[3]
    mid = S.add_matrix(rp, c, v)
    rng = np.random.default_rng(9)
    xs = [rng.uniform(-1, 1, n) for _ in range(9)]
    alone = [S.wait(S.submit(mid, "spmv", x))[0].copy() for x in xs]
    ids = [S.submit(mid, "spmv", x) for x in xs]
    for k, i in enumerate(ids):
        out, _ = S.wait(i)
        assert np.array_equal(out, alone[k]), k
    S.close()


@pytest.mark.parametrize("mode", ["plain", "batched", "mixed"])
def test_large_matrix_tiles_alone_and_in_mix(ctx, pool, mode):
    """A general matrix of 2^20 rows gets column-sorted tile plans and its
    products run one request per tile launch.  The tiles sum each row in
    LDS-atomic order, so alone and in the mix the results agree within the
    stated tolerance (not bit for bit), and each meets the CSR bound against
    the FP64 oracle."""
    mats, lows = pool
    n = 1 << 20
    rp, c, v = gen.zipf_csr(n, 1.5, 256, seed=61)
    S = sp.Server(mode=mode, arena_bytes=256 << 20)
    big = S.add_matrix(rp, c, v)
    small = [S.add_matrix(r, cc, vv) for _, r, cc, vv in mats[:4]]
    rng = np.random.default_rng(62)
    x = rng.uniform(-1, 1, n)
    alone, _ = S.wait(S.submit(big, "spmv", x))
    alone = alone.copy()
    ids = [S.submit(small[k % 4], "spmv", rng.uniform(-1, 1, mats[k % 4][0])) for k in range(20)]
    ids += [S.submit(big, "spmv", rng.uniform(-1, 1, n)) for _ in range(3)]
    probe = S.submit(big, "spmv", x)
    ids += [S.submit(small[k % 4], "spmv", rng.uniform(-1, 1, mats[k % 4][0])) for k in range(20)]
    mixed_out, _ = S.wait(probe)
    yref, _ = oracle.spmv_csr_fp64(n, rp, c, v, x)
    dt = np.float32 if mode == "mixed" else np.float64
    b = spmv_bound(rp, c, v, x, dt)
    assert np.all(np.abs(alone - yref) <= b)
    assert np.all(np.abs(mixed_out - yref) <= b)
    assert np.all(np.abs(mixed_out - alone) <= 2 * b)
    for i in ids:
        S.wait(i)
    S.close()
