"""Autotuner for the SpMV schedule, keyed on row-length statistics.

For a CSR matrix it measures a small set of candidate schedules on the GPU
(HIP-event timing on the context stream) and caches the winner under a key
built from the matrix's row-length statistics (log2 buckets of N, of the
mean and of the maximum row length, and a coefficient-of-variation bucket),
so later matrices of the same class reuse the choice without measuring:

* storage format: CSR, or ELL when the padded width is within the
  reference's 5x guard and cheaper;
* rows per wavefront: the block weight (nnz + row_weight * rows per
  256-thread workgroup) and row weight, which set how many rows share a
  wavefront in the reduction (a block of R rows gets 256/R lanes per row),
  or, for short regular rows, the row-vector kernel's lanes per row;
* bin thresholds: the LDS capacity of a block (twice the block weight)
  decides which rows take the whole-workgroup long-row path;
* column panels for matrices whose x does not fit one XCD's L2;
* column-sorted tiles (allow_tiles): 16384-row tiles over 8192-column
  panels when x does not fit one XCD's L2 (and over 16384-column panels
  when a tile holds fewer than 1024 entries per 8192-column panel), and
  single-panel tiles of up to 4096 rows otherwise.  Tiles sum each row in LDS-atomic order, so a matrix
  tuned to them gives results within the CSR bound but not bit-identical
  from run to run; allow_tiles=False keeps every candidate deterministic.

`plan(A)` returns the chosen schedule built for A, with run(x, y, beta).
"""
import json
import math

import numpy as np

from .core import DeviceArray, sparse_vec_t, create_vec
from .spmv import CsrPlan, CsrPanels, CsrTiles, spmv
from . import convert


def row_stats_key(row_ptr):
    lens = np.diff(np.asarray(row_ptr, np.int64))
    N = len(lens)
    mean = float(lens.mean()) if N else 0.0
    mx = int(lens.max()) if N else 0
    cv = float(lens.std() / mean) if mean > 0 else 0.0
    b = lambda v: int(math.floor(math.log2(max(v, 1))))
    return "N%d_mean%d_max%d_cv%d" % (b(N), b(mean), b(mx), min(int(cv * 2), 8))


class Autotuner:
    BLOCK_WEIGHTS = (256, 512, 1024, 2048)
    ROW_WEIGHTS = (1, 2, 4)

    def __init__(self, ctx, reps=20, allow_tiles=True):
        self.ctx = ctx
        self.reps = reps
        self.allow_tiles = allow_tiles
        self.cache = {}

    def _time(self, run):
        run()
        self.ctx.sync()
        e0, e1 = self.ctx.event(), self.ctx.event()
        self.ctx.record(e0)
        for _ in range(self.reps):
            run()
        self.ctx.record(e1)
        self.ctx.sync()
        return self.ctx.elapsed_ms(e0, e1) / self.reps

    def candidates(self, A, row_ptr):
        N = A.N
        out = []
        for bw in self.BLOCK_WEIGHTS:
            for rw in self.ROW_WEIGHTS:
                out.append({"format": "csr", "block_weight": bw, "row_weight": rw})
        if N * A.dtype.itemsize > (2 << 20):
            out.append({"format": "panels", "block_weight": 1024, "row_weight": 2})
            if self.allow_tiles:
                out.append({"format": "tiles", "tile_rows": 16384, "panel_cols": 8192})
                # few entries per (tile, panel): twice-wider panels as well
                # (the same rule as the C autotuner, csrc/auto.hip)
                nnz = int(row_ptr[-1]) if N else 0
                if nnz / math.ceil(N / 16384) * (8192 / max(N, 1)) < 1024:
                    out.append({"format": "tiles", "tile_rows": 16384, "panel_cols": 16384})
        elif self.allow_tiles and N > 0:
            pc = 64
            while pc < N:
                pc *= 2
            out.append({"format": "tiles", "tile_rows": min(4096, (N + 63) // 64 * 64), "panel_cols": pc})
        lens = np.diff(np.asarray(row_ptr, np.int64))
        if len(lens):
            # row-vector kernel for short, regular rows: lanes around the mean
            # row length, as long as the longest row needs few passes
            mean, mx = float(lens.mean()), int(lens.max())
            g0 = 1 << max(1, int(math.ceil(math.log2(max(mean, 2.0)))))
            for g in (g0 // 2, g0, 2 * g0):
                if 2 <= g <= 64 and mx <= 8 * g:
                    out.append({"format": "csr", "vector_lanes": g})
        if len(lens) and lens.max() * N <= 5 * max(A.nnz, 1):
            out.append({"format": "ell"})
        return out

    def tune(self, A, row_ptr=None):
        """Best schedule for CSR matrix A (host row_ptr optional)."""
        if row_ptr is None:
            row_ptr = A.row.numpy()
        key = row_stats_key(row_ptr) + "_" + A.dtype.name
        if key in self.cache:
            return key, self.cache[key]
        x = DeviceArray(A.N, A.dtype, self.ctx)
        x.fill(1.0)
        y = DeviceArray(A.N, A.dtype, self.ctx)
        best, best_ms, table = None, float("inf"), []
        for cand in self.candidates(A, row_ptr):
            runner = self.build(A, cand)
            ms = self._time(lambda: runner.run(x, y, beta=0.0))
            del runner
            table.append(dict(cand, ms=ms))
            if ms < best_ms:
                best, best_ms = cand, ms
        result = dict(best, ms=best_ms, table=table)
        self.cache[key] = result
        return key, result

    def build(self, A, cand):
        """The schedule `cand` (a tuning result or candidate) built for A:
        an object whose run(x, y, beta) computes y = A.x + beta*y on
        DeviceArrays x and y."""
        fmt = cand["format"]
        if fmt == "csr":
            return _Run(CsrPlan(A, {k: cand[k] for k in ("block_weight", "row_weight", "vector_lanes")
                                    if k in cand}), A.N, vectors=False)
        if fmt == "panels":
            return _Run(CsrPanels(A, 0, {"block_weight": cand["block_weight"], "row_weight": cand["row_weight"]}),
                        A.N, vectors=False)
        if fmt == "tiles":
            return _Run(CsrTiles(A, cand["tile_rows"], cand["panel_cols"]), A.N, vectors=False)
        if fmt == "ell":
            E = convert.csr_ell(A, guard=False)
            return _Run(lambda xv, yv, beta: spmv(E, xv, yv, beta=beta), A.N, vectors=True, keep=E)
        raise ValueError("unknown format %r" % (fmt,))

    def plan(self, A, row_ptr=None):
        """Tune A (or reuse its class's choice) and build that schedule."""
        _, best = self.tune(A, row_ptr)
        return self.build(A, best)

    def save(self, path):
        with open(path, "w") as f:
            json.dump(self.cache, f, indent=1)

    def load(self, path):
        with open(path) as f:
            self.cache.update(json.load(f))


class _Run:
    """Uniform run(x, y, beta) on DeviceArrays x and y over the schedules:
    the plan objects take the arrays, spmv on the ELL copy takes
    sparse_vec_t views of them."""

    def __init__(self, target, N, vectors, keep=None):
        self.target, self.N, self.vectors, self.keep = target, N, vectors, keep

    def run(self, x, y, beta=0.0):
        if not self.vectors:
            return self.target.run(x, y, beta=beta)
        return self.target(sparse_vec_t(x, self.N), sparse_vec_t(y, self.N), beta)


class BatchPlan:
    """Autotuned schedule for a batch of many CSR matrices with nrhs
    right-hand sides each, run as at most two persistent launches.

    The matrices are grouped by row-length class (row_stats_key, with the
    dtype and nrhs); for each class the candidates are measured on (up to
    `sample` of) the class's own matrices in batched launches, and the
    fastest is kept for every matrix of the class (and cached by key):

    * row-block items of one sable_batch_spmv launch, with the block weight
      and row weight of their plans (rows per wavefront: a block packs rows
      until nnz + row_weight * rows reaches the block weight, so a class of
      short rows gets many rows per wave and long rows whole workgroups);
    * column-sorted tile items of one sable_batch_spmv_tiles launch, with
      row tiles of 1024 to 8192 rows over a single column panel (each row
      tile gathers x in ascending column order).

    The per-class assignment is then measured on the whole batch beside the
    uniform schedules (every matrix on row blocks; every matrix on tiles of
    each height), since two launches of parts of a batch do not overlap and
    a class timed alone sees a smaller launch than the batch; the fastest
    whole-batch assignment is kept (`assignment` in info_batch()).

    Row-block items sum each row in a fixed order; tile items in LDS-atomic
    order, so with allow_tiles the results meet the CSR bound rather than a
    fixed bit pattern."""

    ROWBLOCK = ((256, 1), (512, 1), (1024, 1), (2048, 1), (1024, 2), (2048, 2))
    TILE_ROWS = (1024, 2048, 4096, 8192)

    def __init__(self, mats, nrhs=1, reps=5, allow_tiles=True, sample=32, cache=None, row_ptrs=None):
        self.mats, self.nrhs = list(mats), int(nrhs)
        self.ctx = self.mats[0].row.ctx
        self.reps, self.allow_tiles, self.sample = reps, allow_tiles, sample
        self.cache = cache if cache is not None else {}
        rps = row_ptrs if row_ptrs is not None else [A.row.numpy() for A in self.mats]
        self.classes = {}
        for k, (A, rp) in enumerate(zip(self.mats, rps)):
            if A.N == 0:
                continue
            key = "%s_%s_r%d" % (row_stats_key(rp), np.dtype(A.dtype).name, self.nrhs)
            self.classes.setdefault(key, []).append(k)
        for key, idx in self.classes.items():
            if key not in self.cache:
                self.cache[key] = self._tune_class([self.mats[k] for k in idx[:sample]])
        per_class = {}
        for key, idx in self.classes.items():
            for k in idx:
                per_class[k] = self._build(self.mats[k], self.cache[key])
        # whole-batch assignments: per class, and the uniform ones
        uniform = [("all_rowblock", {"format": "rowblock", "block_weight": 0, "row_weight": 0})]
        if allow_tiles:
            uniform += [("all_tiles%d" % tr, {"format": "tiles", "tile_rows": tr})
                        for tr in self.TILE_ROWS if tr * self.nrhs <= 20480]
        xs = [DeviceArray(max(A.N, 1) * self.nrhs, A.dtype, self.ctx) for A in self.mats]
        ys = [DeviceArray(max(A.N, 1) * self.nrhs, A.dtype, self.ctx) for A in self.mats]
        for x in xs:
            x.fill(0.5)
        self.batch_table = []
        best = ("per_class", self._time_batch(self.mats, per_class, xs, ys), per_class)
        self.batch_table.append({"assignment": best[0], "ms": best[1]})
        for name, cand in uniform:
            plans = {k: self._build(self.mats[k], cand) for idx in self.classes.values() for k in idx}
            ms = self._time_batch(self.mats, plans, xs, ys)
            self.batch_table.append({"assignment": name, "ms": ms})
            if ms < best[1]:
                best = (name, ms, plans)
            del plans
        self.assignment, self.batch_ms, self.plans = best

    def _time_batch(self, mats, plans, xs, ys):
        runner = _BatchRunner(mats, plans, xs, ys, self.nrhs, 0.0)
        runner.run()
        self.ctx.sync()
        e0, e1 = self.ctx.event(), self.ctx.event()
        self.ctx.record(e0)
        for _ in range(self.reps):
            runner.run()
        self.ctx.record(e1)
        self.ctx.sync()
        return self.ctx.elapsed_ms(e0, e1) / self.reps

    def info_batch(self):
        """The whole-batch assignment kept, its time, and every time measured."""
        return {"assignment": self.assignment, "ms": self.batch_ms, "table": self.batch_table}

    def _candidates(self):
        out = [{"format": "rowblock", "block_weight": bw, "row_weight": rw} for bw, rw in self.ROWBLOCK]
        if self.allow_tiles:
            out += [{"format": "tiles", "tile_rows": tr} for tr in self.TILE_ROWS if tr * self.nrhs <= 20480]
        return out

    @staticmethod
    def _build(A, cand):
        if cand["format"] == "tiles":
            tr = min(int(cand["tile_rows"]), (A.N + 63) // 64 * 64)
            pc = 64
            while pc < A.N:
                pc *= 2
            return CsrTiles(A, tile_rows=tr, panel_cols=pc)
        if not cand["block_weight"]:
            return A.plan()
        return CsrPlan(A, {"block_weight": cand["block_weight"], "row_weight": cand["row_weight"]})

    def _tune_class(self, mats):
        xs = [DeviceArray(A.N * self.nrhs, A.dtype, self.ctx) for A in mats]
        ys = [DeviceArray(A.N * self.nrhs, A.dtype, self.ctx) for A in mats]
        for x in xs:
            x.fill(0.5)
        best, best_ms, table = None, float("inf"), []
        for cand in self._candidates():
            plans = {k: self._build(A, cand) for k, A in enumerate(mats)}
            ms = self._time_batch(mats, plans, xs, ys)
            table.append(dict(cand, ms=ms))
            if ms < best_ms:
                best, best_ms = cand, ms
            del plans
        return dict(best, ms=best_ms, matrices_timed=len(mats), table=table)

    def info(self):
        """Per class: matrices, the chosen schedule and its time on the
        class's timed sample."""
        return {key: {"matrices": len(idx), **{k: v for k, v in self.cache[key].items() if k != "table"}}
                for key, idx in self.classes.items()}

    def bind(self, xs, ys, beta=0.0):
        """A runner for ys[k] = A_k.xs[k] (+ beta*ys[k]): its item lists
        are on the device, so run() only launches."""
        return _BatchRunner(self.mats, self.plans, xs, ys, self.nrhs, beta)

    def run(self, xs, ys, beta=0.0):
        self.bind(xs, ys, beta).run()


class _BatchRunner:
    """The two launches of a batch: row-block items (sable_batch_spmv) and
    tile items (sable_batch_spmv_tiles), item lists uploaded once."""

    def __init__(self, mats, plans, xs, ys, nrhs, beta):
        from ._lib import call, c_vp, c_i64
        from .core import dtype_code
        from .serve import BatchSpmvItem, _upload_items, _check_vectors, batch_tiles_items
        import ctypes
        self.ctx, self.nrhs, self._call = mats[0].row.ctx, int(nrhs), call
        rb, tl, txs, tys, total = [], [], [], [], 0
        for k, A in enumerate(mats):
            p = plans.get(k)
            if p is None:
                continue
            if isinstance(p, CsrTiles):
                tl.append(p)
                txs.append(xs[k])
                tys.append(ys[k])
                continue
            _check_vectors(A, xs[k], ys[k], nrhs, "BatchPlan")
            blk, nb = c_vp(), c_i64()
            call("sable_csr_plan_blocks", p.handle, ctypes.byref(blk), ctypes.byref(nb))
            rb.append(BatchSpmvItem(A.row.ptr, A.col.ptr, A.val.ptr, blk.value, xs[k].ptr, ys[k].ptr, total,
                                    float(beta)))
            total += nb.value
        self.code = dtype_code(mats[0].dtype)
        self.rb = (_upload_items(rb, self.ctx), len(rb), total) if rb else None
        self.tl = batch_tiles_items(tl, txs, tys, beta, nrhs) if tl else None
        self.keep = (plans, xs, ys)

    def run(self):
        if self.rb is not None:
            d, n, total = self.rb
            self._call("sable_batch_spmv", self.ctx.handle, self.code, self.nrhs, d.ptr, n, total, 0)
        if self.tl is not None:
            d, n, total, maxr, dt, _ = self.tl
            self._call("sable_batch_spmv_tiles", self.ctx.handle, self.code, self.nrhs, d.ptr, n, total, maxr, 0)
