"""Batched launches and the request server (host side).

``batch_spmv`` / ``batch_spts`` / ``batch_spts_refined`` run many matrices
in one persistent launch; ``Server`` wraps the native request server
(continuous batching over resident matrices, HBM-budget admission,
p50/p99 latency).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import call, c_vp, c_i32, c_i64, c_dbl
from .core import Context, DeviceArray, dtype_code


class BatchSpmvItem(ctypes.Structure):
    _fields_ = [("row_ptr", c_vp), ("col", c_vp), ("val", c_vp), ("blk", c_vp), ("x", c_vp), ("y", c_vp),
                ("blk_begin", c_i64), ("beta", c_dbl)]


class BatchSptsItem(ctypes.Structure):
    _fields_ = [("row_ptr", c_vp), ("col", c_vp), ("val", c_vp), ("b", c_vp), ("x", c_vp),
                ("row_begin", c_i64), ("N", c_i32), ("pad", c_i32)]


class BatchRefineItem(ctypes.Structure):
    _fields_ = [("row_ptr", c_vp), ("col", c_vp), ("val64", c_vp), ("val32", c_vp), ("b64", c_vp), ("x64", c_vp),
                ("r32", c_vp), ("d32", c_vp), ("row_begin", c_i64), ("N", c_i32), ("pad", c_i32)]


def _upload_items(items, ctx):
    raw = bytes(memoryview((type(items[0]) * len(items))(*items)).cast("B"))
    return DeviceArray.from_host(np.frombuffer(raw, np.uint8), ctx=ctx)


def batch_spmv(mats, xs, ys, nrhs=1, beta=0.0, workgroups=0, plans=None):
    """ys[k] = mats[k].x[k] (+ beta*ys[k]) for every k in one launch.  x and y
    are DeviceArrays of N*nrhs elements, row-major [N][nrhs].  plans[k]
    (CsrPlan) gives matrix k's row blocks; by default each matrix's own
    plan is used."""
    ctx = mats[0].row.ctx
    items, total = [], 0
    for k, (A, x, y) in enumerate(zip(mats, xs, ys)):
        _check_vectors(A, x, y, nrhs, "batch_spmv")
        blk, nb = c_vp(), c_i64()
        pl = plans[k] if plans is not None else A.plan()
        call("sable_csr_plan_blocks", pl.handle, ctypes.byref(blk), ctypes.byref(nb))
        items.append(BatchSpmvItem(A.row.ptr, A.col.ptr, A.val.ptr, blk.value, x.ptr, y.ptr, total, float(beta)))
        total += nb.value
    d = _upload_items(items, ctx)
    call("sable_batch_spmv", ctx.handle, dtype_code(mats[0].dtype), int(nrhs), d.ptr, len(items), total,
         int(workgroups))
    ctx.sync()


class BatchTilesItem(ctypes.Structure):
    _fields_ = [("tile_ptr", c_vp), ("packed", c_vp), ("val", c_vp), ("x", c_vp), ("y", c_vp),
                ("unit_begin", c_i64), ("beta", c_dbl), ("N", c_i32), ("tile_rows", c_i32), ("ntr", c_i32),
                ("npanels", c_i32), ("panel_cols", c_i32), ("row_bits", c_i32)]


def batch_tile_plans(mats, tile_rows=0, nrhs=1):
    """One column-sorted tile plan per matrix for batch_spmv_tiles: row tiles
    of up to `tile_rows` rows and a single column panel (the next power of
    two >= N), so that each row tile gathers its x in ascending column
    order.  tile_rows * nrhs of the later products must not exceed 20480;
    tile_rows = 0 picks min(4096, 8192 / nrhs), the fastest measured (64 KiB
    of FP64 row sums or less, so two or more workgroups share a CU).
    Matrices with N = 0 get None."""
    from .spmv import CsrTiles
    if not tile_rows:
        tile_rows = min(4096, 8192 // int(nrhs))
    plans = []
    for A in mats:
        if A.N == 0:
            plans.append(None)
            continue
        tr = min(int(tile_rows), (A.N + 63) // 64 * 64)
        pc = 64
        while pc < A.N:
            pc *= 2
        plans.append(CsrTiles(A, tile_rows=tr, panel_cols=pc))
    return plans


def _check_vectors(A, x, y, nrhs, what):
    """x and y must hold N*nrhs values of A's dtype: the kernels index them
    as row-major [N][nrhs] without bounds checks."""
    need = A.N * int(nrhs)
    for v, name in ((x, "x"), (y, "y")):
        if v.dtype != np.dtype(A.dtype) or v.n < need:
            raise ValueError("%s: %s holds %d %s values, the product needs %d %s" %
                             (what, name, v.n, v.dtype.name, need, np.dtype(A.dtype).name))


def batch_tiles_items(plans, xs, ys, beta=0.0, nrhs=1):
    """The device item list of batch_spmv_tiles: (items, nitems, total_units,
    max_tile_rows, dtype, ctx), or None when every plan is None."""
    if not (len(plans) == len(xs) == len(ys)):
        raise ValueError("batch_spmv_tiles: plans, xs and ys differ in length")
    items, total, maxr, ctx, dt = [], 0, 0, None, None
    for p, x, y in zip(plans, xs, ys):
        if p is None:
            continue
        _check_vectors(p.A, x, y, nrhs, "batch_spmv_tiles")
        if dt is None:
            ctx, dt = p.ctx, np.dtype(p.A.dtype)
        elif np.dtype(p.A.dtype) != dt:
            raise ValueError("batch_spmv_tiles: all plans must share one dtype")
        it = BatchTilesItem()
        call("sable_csr_tiles_batch_item", p.handle, x.ptr, y.ptr, float(beta), total, ctypes.byref(it))
        items.append(it)
        total += it.ntr
        maxr = max(maxr, it.tile_rows)
    if not items:
        return None
    return _upload_items(items, ctx), len(items), total, maxr, dt, ctx


def batch_spmv_tiles(plans, xs, ys, nrhs=1, beta=0.0, workgroups=0):
    """ys[k] = A_k.xs[k] (+ beta*ys[k]) for every tile plan (CsrTiles) in one
    persistent launch; x and y are DeviceArrays of N*nrhs elements,
    row-major [N][nrhs] (nrhs in 1, 2, 4, 8; the plans' tile_rows * nrhs
    must not exceed 20480).  None plans (N = 0) are skipped."""
    b = batch_tiles_items(plans, xs, ys, beta, nrhs)
    if b is None:
        return
    d, n, total, maxr, dt, ctx = b
    call("sable_batch_spmv_tiles", ctx.handle, dtype_code(dt), int(nrhs), d.ptr, n, total, maxr, int(workgroups))


def batch_spts(mats, bs, xs, lanes=4):
    ctx = mats[0].row.ctx
    items, total = [], 0
    for L, b, x in zip(mats, bs, xs):
        items.append(BatchSptsItem(L.row.ptr, L.col.ptr, L.val.ptr, b.ptr, x.ptr, total, L.N, 0))
        total += L.N
    d = _upload_items(items, ctx)
    call("sable_batch_spts", ctx.handle, dtype_code(mats[0].dtype), int(lanes), d.ptr, len(items), total)


class BatchLevelPlan:
    """The level schedule of many lower matrices (sable_batch_level_plan_*):
    bind(bs, xs).run() solves every L_k x_k = b_k level by level, one launch
    per level of the deepest matrix, each item with the reference level
    solve's bits."""

    def __init__(self, mats, lanes=4):
        self.mats = list(mats)
        self.ctx = self.mats[0].row.ctx
        items = (BatchSptsItem * len(self.mats))()
        for k, L in enumerate(self.mats):
            items[k] = BatchSptsItem(L.row.ptr, L.col.ptr, L.val.ptr, None, None, 0, L.N, 0)
        h = c_vp()
        call("sable_batch_level_plan_create", self.ctx.handle, int(lanes), ctypes.cast(items, c_vp), len(self.mats),
             ctypes.byref(h))
        self.handle = h

    def info(self):
        lv, ch = c_i32(), c_i64()
        call("sable_batch_level_plan_info", self.handle, ctypes.byref(lv), ctypes.byref(ch))
        return {"levels": lv.value, "chunks": ch.value}

    def bind(self, bs, xs):
        items, total = [], 0
        for L, b, x in zip(self.mats, bs, xs):
            items.append(BatchSptsItem(L.row.ptr, L.col.ptr, L.val.ptr, b.ptr, x.ptr, total, L.N, 0))
            total += L.N
        d = _upload_items(items, self.ctx)
        code = dtype_code(self.mats[0].dtype)

        class _Run:
            def run(_self):
                call("sable_batch_level_plan_solve", self.ctx.handle, self.handle, code, d.ptr, len(items))
        r = _Run()
        r.keep = d
        return r

    def solve(self, bs, xs):
        self.bind(bs, xs).run()

    def bind_refined(self, bs, xs, vals32=None):
        """The FP32 solve with one FP64 refinement step fused into the same
        level sweep, for FP64 matrices: x_k (FP64) from b_k (FP64).  vals32[k]
        (optional) is matrix k's values rounded to FP32; the work space is
        allocated here and kept by the returned runner."""
        items, keep, total = [], [], 0
        for k, (L, b, x) in enumerate(zip(self.mats, bs, xs)):
            v32 = vals32[k] if vals32 is not None else DeviceArray.from_host(
                L.val.numpy().astype(np.float32), ctx=self.ctx)
            w = DeviceArray(2 * L.N, np.float32, self.ctx)
            keep += [v32, w]
            items.append(BatchRefineItem(L.row.ptr, L.col.ptr, L.val.ptr, v32.ptr, b.ptr, x.ptr, w.ptr, None, total,
                                         L.N, 0))
            total += L.N
        d = _upload_items(items, self.ctx)

        class _Run:
            def run(_self):
                call("sable_batch_level_plan_solve_refined", self.ctx.handle, self.handle, d.ptr, len(items))
        r = _Run()
        r.keep = (d, keep)
        return r

    def __del__(self):
        try:
            if self.handle:
                _lib.load().sable_batch_level_plan_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def batch_spts_refined(mats, bs, xs, lanes=4, iters=2):
    """FP64 matrices; FP32 solves with FP64 refinement, all in one batch."""
    ctx = mats[0].row.ctx
    items, keep, total = [], [], 0
    for L, b, x in zip(mats, bs, xs):
        v32 = DeviceArray.from_host(L.val.numpy().astype(np.float32), ctx=ctx)
        r32 = DeviceArray(L.N, np.float32, ctx)
        d32 = DeviceArray(L.N, np.float32, ctx)
        keep += [v32, r32, d32]
        items.append(BatchRefineItem(L.row.ptr, L.col.ptr, L.val.ptr, v32.ptr, b.ptr, x.ptr, r32.ptr, d32.ptr, total,
                                     L.N, 0))
        total += L.N
    d = _upload_items(items, ctx)
    call("sable_batch_spts_refined", ctx.handle, int(lanes), d.ptr, len(items), total, int(iters))


MODES = {"plain": 0, "batched": 1, "mixed": 2}


class Server:
    """The native request server (sable_server_*)."""

    SWEEPS = {"levels": 0, "syncfree": 1}

    def __init__(self, mode="mixed", hbm_budget=0, max_batch=256, slots=3, arena_bytes=256 << 20, refine_iters=1,
                 spts_lanes=4, group_rhs=8, batch_tiles=False, spts_sweep="levels", ctx=None):
        self.ctx = ctx if ctx is not None else Context.default()
        cfg = _lib.ServerConfig(int(hbm_budget), MODES[mode], int(max_batch), int(slots), int(arena_bytes),
                                int(refine_iters), int(spts_lanes), int(group_rhs), int(bool(batch_tiles)),
                                self.SWEEPS[spts_sweep])
        h = c_vp()
        call("sable_server_create", self.ctx.handle, ctypes.byref(cfg), ctypes.byref(h))
        self.handle = h
        self.mode = mode
        self.sizes = []
        self._inflight = {}

    def add_matrix(self, row_ptr, col, val, lower=False):
        row_ptr = np.ascontiguousarray(row_ptr, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        val = np.ascontiguousarray(val, np.float64)
        mid = c_i32()
        call("sable_server_add_matrix", self.handle, len(row_ptr) - 1, row_ptr.ctypes.data, col.ctypes.data,
             val.ctypes.data, 1 if lower else 0, ctypes.byref(mid))
        self.sizes.append(len(row_ptr) - 1)
        return mid.value

    def submit(self, mid, op, vec):
        vec = np.ascontiguousarray(vec, np.float64)
        out = np.zeros(self.sizes[mid], np.float64)
        rid = c_i64()
        call("sable_server_submit", self.handle, int(mid), 0 if op == "spmv" else 1, vec.ctypes.data,
             out.ctypes.data, ctypes.byref(rid))
        self._inflight[rid.value] = (vec, out)
        return rid.value

    def wait(self, rid):
        lat = c_dbl()
        try:
            call("sable_server_wait", self.handle, int(rid), ctypes.byref(lat))
        finally:
            vec, out = self._inflight.pop(rid, (None, None))
        return out, lat.value

    def run_load(self, rate, duration_s, spts_fraction=0.3, seed=1, max_requests=0):
        spec = _lib.LoadSpec(float(rate), float(duration_s), float(spts_fraction), int(seed), int(max_requests))
        st = _lib.LoadStats()
        call("sable_server_run_load", self.handle, ctypes.byref(spec), ctypes.byref(st))
        return {f: getattr(st, f) for f, _ in _lib.LoadStats._fields_}

    def info(self):
        used, n = c_i64(), c_i32()
        call("sable_server_info", self.handle, ctypes.byref(used), ctypes.byref(n))
        return {"hbm_used": used.value, "matrices": n.value}

    def close(self):
        if self.handle:
            _lib.load().sable_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lower_dominant(row_ptr, col, val):
    """Lower triangle with the diagonal last and diagonal = 1 + sum |off|
    (the diagonally dominant lower triangle of a square CSR)."""
    N = len(row_ptr) - 1
    rows = np.repeat(np.arange(N), np.diff(row_ptr))
    keep = col < rows
    r, c, v = rows[keep], col[keep], val[keep]
    cnt = np.bincount(r, minlength=N) + 1
    rp = np.zeros(N + 1, np.int64)
    rp[1:] = np.cumsum(cnt)
    out_c = np.empty(rp[-1], np.int32)
    out_v = np.empty(rp[-1], np.float64)
    # rows are contiguous in r (sorted), so entries keep their order
    starts = np.zeros(N + 1, np.int64)
    starts[1:] = np.cumsum(cnt - 1)
    idx = rp[r] + (np.arange(len(r)) - starts[r])
    out_c[idx] = c
    out_v[idx] = v
    absum = np.bincount(r, weights=np.abs(v), minlength=N)
    out_c[rp[1:] - 1] = np.arange(N)
    out_v[rp[1:] - 1] = 1.0 + absum
    return rp.astype(np.int32), out_c, out_v
