"""Lower sparse triangular solve, mirroring the reference's SpTS surface.

* ``CSR_create_level_sets_with_reorder(A)`` -> a level-ordered CSR carrying
  ``nlevels``, ``level`` (level_ptr, host) and ``permutation`` (device), as
  par/src/sparse32.js:3010-3181 returns.
* ``spts_level_csr(A_level, b_level, y_level)``: the reference-shaped
  level-set solve (par/src/sparse_opt_32.wat:11759-11984).
* ``spts_syncfree(L, b, x)``: the sync-free solve
  (par/src/spts_opt_32.wat:8-260).
* ``SptsAnalysis``: a device-resident analysis for repeated solves (level
  schedule replayed from a hipGraph, sync-free, FP32 + FP64 refinement),
  of a CSR or of a CSC input.
* ``spts_csc(A_csc, x, y, inner_max)``: the serial tree's column sweep
  (seq/src/sparse_opt_32.wat:11-173).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import call, c_vp, c_i32
from .core import DeviceArray, sparse_CSC_t, sparse_CSR_t, dtype_code


class SptsAnalysis:
    def __init__(self, L, lanes=0):
        self.L = L
        self.ctx = L.row.ctx
        h = c_vp()
        if isinstance(L, sparse_CSC_t):
            # the analysis keeps its own CSR copy of the CSC input
            call("sable_spts_csc_analyze", self.ctx.handle, dtype_code(L.dtype), L.N, L.col.ptr, L.row.ptr,
                 L.val.ptr, int(lanes), ctypes.byref(h))
        else:
            call("sable_spts_analyze", self.ctx.handle, dtype_code(L.dtype), L.N, L.row.ptr, L.col.ptr, L.val.ptr,
                 int(lanes), ctypes.byref(h))
        self.handle = h
        nl, mw, ln = c_i32(), c_i32(), c_i32()
        call("sable_spts_info", h, ctypes.byref(nl), ctypes.byref(mw), ctypes.byref(ln))
        self.nlevels, self.max_width, self.lanes = nl.value, mw.value, ln.value

    def set_lanes(self, lanes):
        call("sable_spts_set_lanes", self.handle, int(lanes))
        self.lanes = int(lanes)

    SCHEDULES = {"blocks": 0, "graph": 1, "launches": 2}

    def set_schedule(self, schedule="blocks", block_rows=0):
        """'blocks' (row blocks stepping through the levels, one launch; block_rows a
        power of two, 0 keeps the current), 'graph' (one launch per level in
        a hipGraph) or 'launches' (the per-level launches without a graph)."""
        call("sable_spts_set_schedule", self.handle, self.SCHEDULES[schedule], int(block_rows))

    def tune(self):
        """Times the solve under each schedule and lane count and keeps the
        fastest (the solution is the same bit for bit); returns
        schedule_info() plus the lanes kept and the best time in ms."""
        s, b, ln, ms = ctypes.c_int(), c_i32(), ctypes.c_int(), ctypes.c_float()
        call("sable_spts_tune", self.ctx.handle, self.handle, ctypes.byref(s), ctypes.byref(b), ctypes.byref(ln),
             ctypes.byref(ms))
        self.lanes = ln.value
        info = self.schedule_info()
        info["lanes"] = self.lanes
        info["ms"] = ms.value
        return info

    def schedule_info(self):
        s, b, nb, ns, sl = ctypes.c_int(), c_i32(), c_i32(), c_i32(), ctypes.c_int64()
        call("sable_spts_schedule_info", self.handle, ctypes.byref(s), ctypes.byref(b), ctypes.byref(nb),
             ctypes.byref(ns), ctypes.byref(sl))
        names = {v: k for k, v in self.SCHEDULES.items()}
        return {"schedule": names[s.value], "block_rows": b.value, "nblocks": nb.value, "nsteps": ns.value,
                "slots": sl.value}

    def schedule(self):
        """(level_ptr, permutation, order, row_ptr, col, val) as numpy arrays."""
        L = self.L
        lp = np.zeros(self.nlevels + 1, np.int32)
        perm = DeviceArray(L.N, np.int32, self.ctx)
        order = DeviceArray(L.N, np.int32, self.ctx)
        rp = DeviceArray(L.N + 1, np.int32, self.ctx)
        col = DeviceArray(L.nnz, np.int32, self.ctx)
        val = DeviceArray(L.nnz, L.dtype, self.ctx)
        call("sable_spts_get_schedule", self.ctx.handle, self.handle, lp.ctypes.data, perm.ptr, order.ptr, rp.ptr,
             col.ptr, val.ptr)
        return lp, perm.numpy(), order.numpy(), rp.numpy(), col.numpy(), val.numpy()

    def solve_level(self, b, x):
        call("sable_spts_solve_level", self.ctx.handle, self.handle, b.vec.ptr, x.vec.ptr)

    def solve_syncfree(self, b, x):
        call("sable_spts_solve_syncfree", self.ctx.handle, self.handle, b.vec.ptr, x.vec.ptr)

    def solve_indegree(self, b, x):
        """Sync-free with per-row in-degree counters (same bits as
        solve_syncfree)."""
        call("sable_spts_solve_indegree", self.ctx.handle, self.handle, b.vec.ptr, x.vec.ptr)

    def solve_combined(self, b, x):
        """Level + sync-free combined: the sync-free kernel over the
        level-ordered matrix; equal to solve_level bit for bit."""
        call("sable_spts_solve_combined", self.ctx.handle, self.handle, b.vec.ptr, x.vec.ptr)

    REFINE_INNER = {"level": 0, "syncfree": 1, "fused": 2}

    def solve_refined(self, b, x, iters=2, inner="syncfree"):
        """FP32 solve + `iters` FP64 refinement steps; inner "level" or
        "syncfree" (separate sweeps), or "fused" (iters = 1: the solve and
        its refinement step in one sweep of the schedule)."""
        call("sable_spts_solve_refined", self.ctx.handle, self.handle, b.vec.ptr, x.vec.ptr, int(iters),
             self.REFINE_INNER[inner])

    def status(self):
        call("sable_spts_status", self.ctx.handle, self.handle)

    def __del__(self):
        try:
            if self.handle:
                _lib.load().sable_spts_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def CSR_create_level_sets_with_reorder(A, lanes=0):
    """par/src/sparse32.js:3010-3181: the level-ordered CSR with nlevels,
    level (level_ptr, host int32 array) and permutation (device)."""
    an = SptsAnalysis(A, lanes)
    lp, perm, order, rp, col, val = an.schedule()
    B = sparse_CSR_t(DeviceArray.from_host(rp, ctx=A.row.ctx), DeviceArray.from_host(col, ctx=A.row.ctx),
                     DeviceArray.from_host(val, ctx=A.row.ctx), A.N, A.nnz)
    B.nlevels = an.nlevels
    B.level = lp
    B.permutation = DeviceArray.from_host(perm, ctx=A.row.ctx)
    B.order = order
    B.analysis = an
    return B


def CSR_create_level_sets(A, lanes=0):
    """par/src/sparse32.js:2942-3008: sets A.nlevels, A.level, A.permutation."""
    an = SptsAnalysis(A, lanes)
    lp, perm, order, _, _, _ = an.schedule()
    A.nlevels, A.level, A.permutation = an.nlevels, lp, DeviceArray.from_host(perm, ctx=A.row.ctx)
    A.analysis = an
    return A


def spts_level_csr(A, b_view, y_view, lanes=1):
    """Level-set solve of a level-ordered CSR with level-ordered b and y."""
    call("sable_spts_level_csr", A.row.ctx.handle, dtype_code(A.dtype), A.N, A.nlevels,
         np.ascontiguousarray(A.level, np.int32).ctypes.data, A.row.ptr, A.col.ptr, A.val.ptr, b_view.vec.ptr,
         y_view.vec.ptr, int(lanes))


def spts_syncfree(L, b_view, x_view, lanes=1, spin_limit=0):
    """Sync-free solve (natural order)."""
    call("sable_spts_syncfree_csr", L.row.ctx.handle, dtype_code(L.dtype), L.N, L.row.ptr, L.col.ptr, L.val.ptr,
         b_view.vec.ptr, x_view.vec.ptr, int(lanes), int(spin_limit))


def spts_csc(A, x_view, y_view, inner_max=1):
    """seq/src/sparse_opt_32.wat:11-173 (spts_csc_wrapper): y = L^-1 x for
    a lower CSC with the diagonal first in each column, bit-identical to the
    column sweep."""
    call("sable_spts_csc", A.col.ctx.handle, dtype_code(A.dtype), A.N, A.col.ptr, A.row.ptr, A.val.ptr,
         x_view.vec.ptr, y_view.vec.ptr, int(inner_max))
