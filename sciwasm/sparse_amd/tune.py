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
  panels when x does not fit one XCD's L2, and single-panel tiles of up to
  4096 rows otherwise.  Tiles sum each row in LDS-atomic order, so a matrix
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
