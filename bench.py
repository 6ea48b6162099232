#!/usr/bin/env python3
"""Benchmark of the Level-II hot path on one MI355X.

Headline (the JSON line's `value`): FP64 CSR SpMV y = A.x on a random
4,194,304 x 4,194,304 matrix whose row lengths follow the Zipf law with
s = 1.5 truncated to [1, 4096] (renormalised; mean about 49, about 206M
nonzeros), per-row columns uniform without replacement, values N(0,1),
x ~ U(-1,1), int32 indices, float64 values; inputs resident in HBM.  One
step is one whole product.  Bytes per step (algorithmic) = 4(N+1) +
12 nnz + 8N (x once) + 8N (y).  The fastest of the measured schedules is
reported (row-block CSR, the column-panel blocked plan, or the column-sorted
tile plan of the same CSR).

Also measured and reported under "extra" (each section bounded to seconds):
  c1      FP32 5-point Laplacian 256x256 (65,536 rows): CSR/ELL/DIA/COO
          SpMV in microseconds and GB/s, reference checksum protocol
  addon   the reference's call surface: spmv_csr_wrapper through the
          Node-API addon (node sciwasm/sparse_amd/node/bench_node.js) on the
          c1 matrix (inner_max 1 and 500) and on the headline matrix
          (inner_max 1 and 100), x and y across PCIe on every call
  band    FP64 7-diagonal band, 16,777,216 rows (offsets 0, +-1, +-64,
          +-4096): csr_dia / csr_ell / coo_csr / transpose on the GPU and
          DIA/ELL/CSR SpMV GB/s
  zipf16  the same product on the sparser Zipf(s = 1.73) matrix of the same
          N (mean about 16, about 67M nonzeros), two tile shapes
  convert coo_csr of the headline matrix (shuffled COO, 206M entries); COO
          SpMV of the same matrix, shuffled (atomics) and row-sorted
  ingest  MatrixMarket text parsed on the device and expanded to COO
  ewise   element-wise ops and eliminate_zeros on the headline values
  reorder sort_rows_by_nnz of the headline matrix, reorder_NN of a
          scattered 2D Laplacian
  spts    lower solves, 3D 7-point Laplacian 128^3 and a random 1,048,576-row
          lower matrix: analysis, level-set solve over row blocks (one
          launch) and over levels (hipGraph, or per-level launches with
          --no-spts-graph), the tuned choice, sync-free, combined, FP32 +
          FP64 refinement; residuals and bit-exactness against the oracle
  batch   2,048 matrices in one persistent launch vs one launch per matrix
  batch_tiles  (not in the default list) only the FP64 batched tile product
          of the same pool, three launches, for hardware-counter runs
  serve   open-loop Poisson serving, plain FP64 one launch per request vs
          batched FP64 and batched FP32 + FP64 refinement (also with the
          pool's SpMV as batched tile plans, "mixed+tiles"): solves/s,
          p50/p99 latency
The CPU baseline is the serial C restatement of the reference kernels on
this host (the reference's own Wasm needs a WAT assembler and Node >= 16.4,
which this image does not have).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--sections ...]
"""
import argparse
import json
import os
import sys
import time
import traceback

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# sable_csr_tiles_select modes: (variant name, kernel)
TILE_MODES = {0: ("perpanel", "k_tiles"), 1: ("fused", "k_tiles_fused")}
# Fabric/HBM bytes per product measured with rocprofv3 FETCH_SIZE and
# WRITE_SIZE (one counter per run, scaled by the readings of a 1 GiB copy:
# FETCH x2, WRITE x1) for the headline matrix (seed 20261203) and schedule;
# profiles/csr_spmv_fp64_zipf4M_*.md hold the runs.  Each entry names the
# source file of its kernel and the SHA-256 of that file when measured: a
# changed kernel source reports traffic null until it is measured again.
MEASURED_TRAFFIC = {
    "tiles_16384x8192_fused": {
        "per_product": 2.858e9, "per_launch": 2.858e9,
        "source": "sciwasm/sparse_amd/csrc/spmv_tiles.hip",
        "sha256": "4860447ddc8b9ea8673fab2724c8901a6d783cc793bb54c04ef3da5088313357",
        "limiter": "L2 request rate: TCC busy 97% of the kernel's cycles at 0.47 requests per cycle per "
                   "channel; ~71M x-line requests (every CU touches most of x's lines: 805k entries over "
                   "262k 128-byte lines) beside ~19M streaming requests and 2.1M x prefetches; fetched "
                   "traffic 1.12x the algorithmic bytes"},
}


def measured_traffic(variant, N):
    """The measured traffic of a schedule on the headline matrix, or None
    when it was not measured or its kernel source has changed since."""
    import hashlib
    m = MEASURED_TRAFFIC.get(variant)
    if m is None or N != 4194304 or m["sha256"] is None:
        return None
    with open(os.path.join(ROOT, m["source"]), "rb") as f:
        if hashlib.sha256(f.read()).hexdigest() != m["sha256"]:
            return None
    return m


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Timer:
    def __init__(self, ctx):
        self.ctx = ctx

    def __call__(self, run, steps, warmup):
        for _ in range(warmup):
            run()
        self.ctx.sync()
        e0, e1 = self.ctx.event(), self.ctx.event()
        self.ctx.record(e0)
        for _ in range(steps):
            run()
        self.ctx.record(e1)
        self.ctx.sync()
        return self.ctx.elapsed_ms(e0, e1) / steps

    def once(self, run):
        self.ctx.sync()
        e0, e1 = self.ctx.event(), self.ctx.event()
        self.ctx.record(e0)
        r = run()
        self.ctx.record(e1)
        self.ctx.sync()
        return self.ctx.elapsed_ms(e0, e1), r


def r4(v):
    return round(float(v), 4)


# ------------------------------------------------------------------ headline
def headline(args, sp, ctx, T):
    from sciwasm.sparse_amd import gen, _lib as L
    N = args.rows
    t0 = time.time()
    rp, col, val = gen.zipf_csr(N, args.zipf_s, 4096, seed=args.seed, dtype=np.float64)
    nnz = int(rp[-1])
    x = np.random.default_rng(args.seed + 1).uniform(-1.0, 1.0, N)
    log("generated N=%d nnz=%d (mean %.2f, max %d) in %.1fs" % (N, nnz, nnz / N, np.diff(rp).max(), time.time() - t0))
    A = sp.csr_from_host(N, rp, col, val, ctx)
    xv = sp.vec_from_host(x, ctx=ctx)
    yv = sp.create_vec(N, 0.0, np.float64, ctx=ctx)
    bytes_step = 4 * (N + 1) + 12 * nnz + 8 * N + 8 * N
    if args.calib_copy:
        a = sp.DeviceArray(1 << 27, np.float64, ctx)
        b = sp.DeviceArray(1 << 27, np.float64, ctx)
        a.fill(1.0)
        L.call("sable_memcpy_d2d", ctx.handle, b.ptr, a.ptr, a.nbytes)
        ctx.sync()
        del a, b
    variants, results = {}, {}
    for kern in [int(v) for v in args.csr_kernel.split(",") if v.strip()]:
        L.call("sable_csr_kernel_select", kern)
        if not args.no_rowblock:
            plan = A.plan()
            ms = T(lambda: plan.run(xv.vec, yv.vec, beta=0.0), args.steps, args.warmup)
            key = "rowblock_k%d" % kern
            variants[key] = {"ms": ms, "gbs": bytes_step / (ms * 1e-3) / 1e9, "launches_per_step": 1, "csr_kernel": kern}
            results[key] = yv.array()
            log(key, plan.info(), "%.4f ms  %.1f GB/s" % (ms, variants[key]["gbs"]))
            del plan
            A._plan = None
        for pc in [int(v) for v in args.panel_cols.split(",") if v.strip()]:
            for bw in [int(v) for v in args.block_weight.split(",") if v.strip()]:
                pan = sp.CsrPanels(A, panel_cols=pc, tune={"block_weight": bw} if bw else None)
                info = pan.info()
                ms = T(lambda: pan.run(xv.vec, yv.vec, beta=0.0), args.steps, args.warmup)
                key = "panels_%d_bw%d_k%d" % (info["panel_cols"], bw or 1024, kern)
                variants[key] = {"ms": ms, "gbs": bytes_step / (ms * 1e-3) / 1e9,
                                 "launches_per_step": info["npanels"], "npanels": info["npanels"],
                                 "panel_cols": info["panel_cols"], "block_weight": bw or 1024, "csr_kernel": kern}
                results[key] = yv.array()
                log(key, info, "%.4f ms  %.1f GB/s" % (ms, variants[key]["gbs"]))
                del pan
    L.call("sable_csr_kernel_select", 0)
    for spec in [v for v in args.tiles.split(",") if v.strip()]:
        tr, pc = (int(t) for t in spec.split("x"))
        t1 = time.time()
        tl = sp.CsrTiles(A, tr, pc)
        info = tl.info()
        build_s = time.time() - t1
        for mode in [int(v) for v in args.tiles_mode.split(",") if v.strip()]:
            L.call("sable_csr_tiles_select", mode)
            ms = T(lambda: tl.run(xv.vec, yv.vec, beta=0.0), args.steps, args.warmup)
            key = "tiles_%dx%d_%s" % (tr, pc, TILE_MODES[mode][0])
            variants[key] = {"ms": ms, "gbs": bytes_step / (ms * 1e-3) / 1e9,
                             "launches_per_step": 1 if mode else info["npanels"],
                             "npanels": info["npanels"], "row_tiles": info["row_tiles"], "tile_rows": tr,
                             "panel_cols": pc, "kernel": TILE_MODES[mode][1],
                             "plan_build_s": round(build_s, 3),
                             # the plan's own copy of the matrix, resident beside the caller's CSR
                             "plan_extra_hbm_bytes": info["bytes"]}
            results[key] = yv.array()
            log(key, info, "built in %.2fs, %.4f ms  %.1f GB/s" % (build_s, ms, variants[key]["gbs"]))
        L.call("sable_csr_tiles_select", 1)
        del tl
    best = min(variants, key=lambda k: variants[k]["ms"])
    ms, gbs, y = variants[best]["ms"], variants[best]["gbs"], results[best]
    log("best %s: %.4f ms/step, %.1f GB/s (%.1f%% of HBM peak)" % (best, ms, gbs, 100 * gbs / HBM_PEAK_GBS))
    cpu, check = None, None
    if not args.no_cpu:
        import oracle
        R = N // 4
        nnz_s = int(rp[R])
        t, ycpu = oracle.time_spmv_csr_f64(R, rp, col, val, x, warm=10, reps=60)
        bytes_s = 4 * (R + 1) + 12 * nnz_s + 8 * N + 8 * R
        serial = {"value": round(bytes_s / t / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
                  "sample": "serial C restatement of seq/src/sparse_opt_64.wat spmv_csr on rows [0, N/4) "
                            "(%d rows, %d nnz) of the same matrix, 10 warm-up + 60 timed calls, %.4f s/call"
                            % (R, nnz_s, t)}
        try:
            W = len(os.sched_getaffinity(0))
        except AttributeError:
            W = os.cpu_count() or 1
        W = max(1, min(W, 64))
        tw, yw = oracle.time_spmv_csr_f64_mt(N, rp, col, val, x, W, warm=3, reps=60)
        cpu = {"value": round(bytes_step / tw / 1e9, 3), "unit": "GB/s", "cores": W, "kind": "port",
               "sample": "the same C restatement on %d persistent threads over static_nnz row ranges "
                         "(par/src/sparse32.js:2769-2821), whole matrix (%d rows, %d nnz), 3 warm-up + 60 "
                         "timed calls, %.4f s/call; results identical to the serial restatement: %s"
                         % (W, N, nnz, tw, bool(np.array_equal(yw[:R], ycpu))),
               "serial": serial}
        lens = np.diff(rp[:R + 1]).astype(np.float64) + 1
        ab = np.add.reduceat(np.abs(val[:nnz_s] * x[col[:nnz_s]]), rp[:R])
        ab[np.diff(rp[:R + 1]) == 0] = 0
        u = np.finfo(np.float64).eps / 2
        bnd = 2 * lens * u / (1 - lens * u) * ab
        check = {"rows_checked": R,
                 "max_abs_err_over_bound": float(np.max(np.abs(y[:R] - ycpu) / np.maximum(bnd, 1e-300))),
                 "ok": bool(np.all(np.abs(y[:R] - ycpu) <= bnd + 1e-300))}
        log("cpu baseline:", cpu, "check:", check)
    head = {"N": N, "nnz": nnz, "ms": ms, "gbs": gbs, "best": best, "bytes_step": bytes_step,
            "variants": {k: {kk: (r4(vv) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                         for k, v in variants.items()},
            "cpu": cpu, "check": check}
    host = (rp, col, val, x)
    del A, xv, yv
    return head, host


# ------------------------------------------------------------------ c1
def sec_c1(args, sp, ctx, T):
    import oracle
    from sciwasm.sparse_amd import gen
    n = 256
    rp, col, val = gen.laplacian_2d(n, np.float32)
    N = n * n
    nnz = len(col)
    x = np.random.default_rng(1).uniform(-1, 1, N).astype(np.float32)
    A = sp.csr_from_host(N, rp, col, val, ctx)
    E = sp.csr_ell(A)
    D = sp.csr_dia(A)
    row = gen.csr_to_coo_rows(rp)
    C = sp.coo_from_host(N, row, col, val, ctx)
    xv, yv = sp.vec_from_host(x, ctx=ctx), sp.create_vec(N, 0.0, np.float32, ctx=ctx)
    out = {"N": N, "nnz": nnz}
    fmt_bytes = {"csr": 4 * (N + 1) + 8 * nnz + 8 * N, "ell": 8 * E.ncols * N + 8 * N,
                 "dia": 4 * D.ndiags * N + 4 * D.ndiags + 8 * N, "coo": 12 * nnz + 8 * N}
    for name, M in (("csr", A), ("ell", E), ("dia", D), ("coo", C)):
        ms = T(lambda: sp.spmv(M, xv, yv, beta=0.0), 200, 20)
        out[name] = {"us": r4(ms * 1e3), "gbs": r4(fmt_bytes[name] / (ms * 1e-3) / 1e9)}
    for g in (4, 8):
        pl = sp.CsrPlan(A, {"vector_lanes": g})
        ms = T(lambda: pl.run(xv.vec, yv.vec, beta=0.0), 200, 20)
        out["csr_vec%d" % g] = {"us": r4(ms * 1e3), "gbs": r4(fmt_bytes["csr"] / (ms * 1e-3) / 1e9)}
    # reference protocol (seq/tests/main32.js): x[i] = i, y = 0, y += A.x, checksum
    sp.init_x(xv)
    sp.clear_vec(yv)
    sp.spmv(A, xv, yv)
    ydev = yv.array()
    yref = oracle.spmv_csr(N, rp, col, val, np.arange(N, dtype=np.float32), np.zeros(N, np.float32))
    out["fletcher_gpu"] = oracle.fletcher_sum(ydev)
    out["fletcher_ref_order"] = oracle.fletcher_sum(yref)
    out["max_abs_diff_vs_ref_order"] = float(np.max(np.abs(ydev - yref)))
    # CPU baseline of the same product: serial restatement, timed in C over many calls
    t0 = time.perf_counter()
    reps = 300
    oracle.spmv_csr(N, rp, col, val, x, np.zeros(N, np.float32), reps)
    t = (time.perf_counter() - t0) / reps
    out["cpu_serial_us"] = r4(t * 1e6)
    log("c1:", out)
    return out


# ------------------------------------------------------------------ addon
def _addon_run(d, bits, rp, col, val, x, warmup, calls, inner, linear_mb):
    """Writes the matrix for bench_node.js and runs it in a child Node process
    (the addon holds its own HIP context there)."""
    import shutil
    import subprocess
    N, nnz = len(rp) - 1, len(col)
    for name, a in (("rp", rp), ("col", col), ("val", val), ("x", x)):
        np.ascontiguousarray(a).tofile(os.path.join(d, name + ".bin"))
    with open(os.path.join(d, "spec.json"), "w") as f:
        json.dump({"N": N, "nnz": nnz}, f)
    node = shutil.which("node")
    if node is None:
        raise RuntimeError("node is not installed")
    script = os.path.join(ROOT, "sciwasm", "sparse_amd", "node", "bench_node.js")
    r = subprocess.run([node, script, d, str(bits), str(warmup), str(calls), ",".join(str(i) for i in inner)],
                       capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, SABLE_LINEAR_MEMORY_MB=str(linear_mb)))
    if r.returncode != 0:
        raise RuntimeError("bench_node.js failed: " + (r.stdout[-2000:] + r.stderr[-2000:]))
    out = json.loads(r.stdout.strip().splitlines()[-1])
    y = np.fromfile(os.path.join(d, "y.bin"), dtype=val.dtype)
    return out, y


def _row_check(rp, col, val, x, y, R):
    """y[:R] against the reference-order oracle, within the CSR bound."""
    import oracle
    yref = oracle.spmv_csr(R, rp[:R + 1], col[:rp[R]], val[:rp[R]], x, np.zeros(R, val.dtype), 1)
    lens = np.diff(rp[:R + 1]).astype(np.float64) + 1
    ab = np.add.reduceat(np.abs(val[:rp[R]].astype(np.float64) * x.astype(np.float64)[col[:rp[R]]]), rp[:R])
    ab[np.diff(rp[:R + 1]) == 0] = 0
    u = np.finfo(val.dtype).eps / 2
    bnd = 2 * lens * u / (1 - lens * u) * ab + np.finfo(val.dtype).tiny
    return bool(np.all(np.abs(y[:R].astype(np.float64) - yref.astype(np.float64)) <= bnd))


def sec_addon(args, sp, ctx, host):
    """The reference's call surface, timed: spmv_csr_wrapper through the
    Node-API addon (bench_node.js), matrix resident after the first call,
    x and y resident while the caller leaves them unchanged (as in the
    reference's own timing loop) and y brought back after every call, on C1
    (FP32 Laplacian 256^2, inner_max 1 and 500) and on the headline matrix
    (FP64, inner_max 1 and 100); and spts_csr_sync_free_wrapper on the two
    C3 lower matrices (FP64, b resident, the solution brought back)."""
    import tempfile
    import shutil
    import subprocess
    import oracle
    from sciwasm.sparse_amd import gen
    names = {-1: "none", 0: "row blocks", 1: "row vector", 2: "column-sorted tiles"}
    out = {}
    base = tempfile.mkdtemp(prefix="sable_addon_")
    try:
        rp, col, val = gen.laplacian_2d(256, np.float32)
        N = len(rp) - 1
        x = np.random.default_rng(1).uniform(-1, 1, N).astype(np.float32)
        d = os.path.join(base, "c1")
        os.makedirs(d)
        r, y = _addon_run(d, 32, rp, col, val, x, 20, 200, (1, 500), 64)
        t0 = time.perf_counter()
        reps = 300
        oracle.spmv_csr(N, rp, col, val, x, np.zeros(N, np.float32), reps)
        cpu_us = (time.perf_counter() - t0) / reps * 1e6
        b = 4 * (N + 1) + 8 * len(col) + 8 * N
        out["c1"] = {"N": N, "nnz": len(col), "schedule": names.get(r["schedule"], r["schedule"]),
                     "first_call_ms": r4(r["first_call_ms"]),
                     "call_us_inner1": r4(r["runs"]["1"]["median_ms"] * 1e3),
                     "call_ms_inner500": r4(r["runs"]["500"]["median_ms"]),
                     "product_us_inner500": r4(r["runs"]["500"]["median_ms"] * 1e3 / 500),
                     "gbs_inner500": r4(b * 500 / (r["runs"]["500"]["median_ms"] * 1e-3) / 1e9),
                     "uploads_in_timed_calls": r["runs"]["1"]["uploads"] + r["runs"]["500"]["uploads"],
                     "vector_uploads_in_timed_calls": (r["runs"]["1"]["vector_uploads"] +
                                                       r["runs"]["500"]["vector_uploads"]),
                     "cpu_serial_us": r4(cpu_us), "within_csr_bound": _row_check(rp, col, val, x, y, N),
                     "write_tracking": r["write_tracking"]}
        log("addon c1:", out["c1"])
        if host is not None:
            rp, col, val, x = host
            N, nnz = len(rp) - 1, len(col)
            d = os.path.join(base, "c2a")
            os.makedirs(d)
            lin = int((4 * (N + 1) + 12 * nnz + 16 * N) / (1 << 20)) + 256
            im = 100
            r, y = _addon_run(d, 64, rp, col, val, x, 2, 10, (1, im), lin)
            b = 4 * (N + 1) + 12 * nnz + 16 * N
            c1m, cim = r["runs"]["1"]["median_ms"], r["runs"][str(im)]["median_ms"]
            # y down on every call, plus any x or y upload the timed calls made
            vup = r["runs"]["1"]["vector_uploads"] / r["runs"]["1"]["calls"]
            pcie = 8 * N * (1 + vup)
            # per-call cost = transfers + products: the two call shapes give
            # the transfer part and the product part apart (a long call keeps
            # the GPU busy, so its products run at the clocks of a busy GPU)
            prod = (cim - c1m) / (im - 1)
            out["c2a"] = {"N": N, "nnz": nnz, "schedule": names.get(r["schedule"], r["schedule"]),
                          "load_s": r4(r["load_s"]), "first_call_ms": r4(r["first_call_ms"]),
                          "call_ms_inner1": r4(c1m), "call_ms_inner%d" % im: r4(cim),
                          "product_ms_from_calls": r4(prod), "transfer_ms_per_call": r4(c1m - prod),
                          "pcie_gbs": r4(pcie / ((c1m - prod) * 1e-3) / 1e9) if c1m > prod else None,
                          "gbs_inner1": r4(b / (c1m * 1e-3) / 1e9),
                          "gbs_inner%d" % im: r4(im * b / (cim * 1e-3) / 1e9),
                          "uploads_in_timed_calls": r["runs"]["1"]["uploads"] + r["runs"][str(im)]["uploads"],
                          "vector_uploads_in_timed_calls": (r["runs"]["1"]["vector_uploads"] +
                                                            r["runs"][str(im)]["vector_uploads"]),
                          "rows_checked": N // 4, "within_csr_bound": _row_check(rp, col, val, x, y, N // 4)}
            log("addon c2a:", out["c2a"])
        # the reference's sync-free lower solve (spts_csr_sync_free_wrapper)
        # on the two C3 matrices, FP64, bit-exact against the oracle's
        # sync-free order; CPU: the serial restatement of the same solve
        for name, make in (("c3a_lap3d_128", lambda: gen.laplacian_3d_lower(128)),
                           ("c3b_random_lower_1M", lambda: gen.random_lower(1 << 20, 32, seed=7))):
            Lrp, Lc, Lv = make()
            n = len(Lrp) - 1
            b = np.random.default_rng(3).uniform(-1, 1, n)
            d = os.path.join(base, name)
            os.makedirs(d)
            for fname, arr in (("rp", Lrp), ("col", Lc), ("val", Lv), ("b", b)):
                np.ascontiguousarray(arr).tofile(os.path.join(d, fname + ".bin"))
            with open(os.path.join(d, "spec.json"), "w") as f:
                json.dump({"N": n, "nnz": len(Lc)}, f)
            lin = int((4 * (n + 1) + 12 * len(Lc) + 24 * n) / (1 << 20)) + 128
            r = subprocess.run([shutil.which("node"), os.path.join(ROOT, "sciwasm", "sparse_amd", "node",
                                                                   "bench_node_spts.js"), d, "64", "3", "20"],
                               capture_output=True, text=True, timeout=600,
                               env=dict(os.environ, SABLE_LINEAR_MEMORY_MB=str(lin)))
            if r.returncode != 0:
                raise RuntimeError("bench_node_spts.js failed: " + (r.stdout[-2000:] + r.stderr[-2000:]))
            rs = json.loads(r.stdout.strip().splitlines()[-1])
            y = np.fromfile(os.path.join(d, "y.bin"), dtype=np.float64)
            t0 = time.perf_counter()
            ys = oracle.spts_syncfree(n, Lrp, Lc, Lv, b)
            cpu_ms = (time.perf_counter() - t0) * 1e3
            out[name] = {"N": n, "nnz": len(Lc), "first_call_ms": r4(rs["first_call_ms"]),
                         "call_ms": r4(rs["median_ms"]), "uploads_in_timed_calls": rs["uploads"],
                         "vector_uploads_in_timed_calls": rs["vector_uploads"],
                         "cpu_serial_ms": r4(cpu_ms), "bit_exact_vs_oracle_syncfree": bool(np.array_equal(y, ys))}
            log("addon %s:" % name, out[name])
    finally:
        shutil.rmtree(base, ignore_errors=True)
    return out


# ------------------------------------------------------------------ band
def sec_zipf16(args, sp, ctx, T):
    """The same FP64 CSR SpMV on the sparser Zipf matrix whose size matches
    the ~16 mean / ~67 M nnz quoted beside the literal s = 1.5 law: row
    lengths Zipf(s = 1.73) truncated to [1, 4096]."""
    from sciwasm.sparse_amd import gen
    N = args.rows
    rp, col, val = gen.zipf_csr(N, 1.73, 4096, seed=args.seed, dtype=np.float64)
    nnz = int(rp[-1])
    x = np.random.default_rng(args.seed + 1).uniform(-1.0, 1.0, N)
    A = sp.csr_from_host(N, rp, col, val, ctx)
    xv, yv = sp.vec_from_host(x, ctx=ctx), sp.create_vec(N, 0.0, np.float64, ctx=ctx)
    bytes_step = 4 * (N + 1) + 12 * nnz + 8 * N + 8 * N
    out = {"N": N, "nnz": nnz, "mean_row": round(nnz / N, 3), "zipf_s": 1.73, "bytes_step": bytes_step}
    R = N // 8
    ab = np.abs(val[:rp[R]] * x[col[:rp[R]]])
    yref = np.add.reduceat(val[:rp[R]] * x[col[:rp[R]]], rp[:R])
    sab = np.add.reduceat(ab, rp[:R])
    lens = np.diff(rp[:R + 1])
    yref[lens == 0] = 0
    sab[lens == 0] = 0
    u = np.finfo(np.float64).eps / 2
    bnd = 2 * (lens + 1) * u / (1 - (lens + 1) * u) * sab + 1e-300
    shapes = [tuple(int(t) for t in v.split("x")) for v in args.zipf16_tiles.split(",") if v.strip()]
    for tr, pc in shapes:
        tl = sp.CsrTiles(A, tr, pc)
        ms = T(lambda: tl.run(xv.vec, yv.vec, beta=0.0), args.steps, args.warmup)
        y = yv.array()[:R]
        gbs = bytes_step / (ms * 1e-3) / 1e9
        out["tiles_%dx%d_fused" % (tr, pc)] = {"ms": r4(ms), "gbs": r4(gbs), "pct_hbm": r4(100 * gbs / HBM_PEAK_GBS),
                                              "rows_checked": R,
                                              "within_csr_bound": bool(np.all(np.abs(y - yref) <= bnd))}
        del tl
    # the schedule the reference's spmv_csr gets through the addon
    sp.CsrAuto.clear_cache()
    auto = sp.CsrAuto(A)
    ai = auto.info()
    ms = T(lambda: auto.run(xv.vec, yv.vec, beta=0.0), args.steps, args.warmup)
    y = yv.array()[:R]
    gbs = bytes_step / (ms * 1e-3) / 1e9
    out["autotuned"] = {"schedule": ai["schedule"], "param0": ai["param0"], "param1": ai["param1"], "ms": r4(ms),
                        "gbs": r4(gbs), "pct_hbm": r4(100 * gbs / HBM_PEAK_GBS), "rows_checked": R,
                        "within_csr_bound": bool(np.all(np.abs(y - yref) <= bnd))}
    del auto
    log("zipf16:", out)
    return out


def sec_band(args, sp, ctx, T):
    from sciwasm.sparse_amd import gen
    N = 1 << 24
    offs = [-4096, -64, -1, 0, 1, 64, 4096]
    t0 = time.time()
    rp, col, val = gen.band_csr(N, offs, seed=5)
    nnz = len(col)
    log("band generated nnz=%d in %.1fs" % (nnz, time.time() - t0))
    x = np.random.default_rng(2).uniform(-1, 1, N)
    A = sp.csr_from_host(N, rp, col, val, ctx)
    xv, yv = sp.vec_from_host(x, ctx=ctx), sp.create_vec(N, 0.0, np.float64, ctx=ctx)
    out = {"N": N, "nnz": nnz, "offsets": offs}
    t_dia, D = T.once(lambda: sp.csr_dia(A))
    t_ell, E = T.once(lambda: sp.csr_ell(A))
    out["csr_dia_ms"] = r4(t_dia)
    out["csr_ell_ms"] = r4(t_ell)
    out["dia_guard"] = "accepted" if D is not None else "refused"
    b_dia = 8 * D.ndiags * N + 4 * D.ndiags + 16 * N
    b_ell = 12 * E.ncols * N + 16 * N
    b_csr = 4 * (N + 1) + 12 * nnz + 16 * N
    for name, M, b in (("dia", D, b_dia), ("ell", E, b_ell), ("csr", A, b_csr)):
        ms = T(lambda: sp.spmv(M, xv, yv, beta=0.0), 20, 3)
        out[name + "_spmv"] = {"ms": r4(ms), "gbs": r4(b / (ms * 1e-3) / 1e9),
                               "pct_hbm": r4(100 * b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS)}
    # CSR with the row-vector kernel (lanes per row around the row length 7)
    for g, rr in ((8, 1), (8, 2), (8, 4), (8, 8), (8, 16), (4, 8), (16, 4)):
        pl = sp.CsrPlan(A, {"vector_lanes": g, "vector_rows": rr})
        ms = T(lambda: pl.run(xv.vec, yv.vec, beta=0.0), 20, 3)
        out["csr_vec%dx%d_spmv" % (g, rr)] = {"ms": r4(ms), "gbs": r4(b_csr / (ms * 1e-3) / 1e9),
                                              "pct_hbm": r4(100 * b_csr / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS)}
        del pl
    # CSR through column-sorted tiles (16384-row tiles over 8192-column panels)
    yc = yv.array()
    tl = sp.CsrTiles(A)
    ms = T(lambda: tl.run(xv.vec, yv.vec, beta=0.0), 20, 3)
    out["csr_tiles_spmv"] = {"ms": r4(ms), "gbs": r4(b_csr / (ms * 1e-3) / 1e9),
                             "pct_hbm": r4(100 * b_csr / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS),
                             "max_rel_diff_vs_csr": float(np.max(np.abs(yv.array() - yc)) /
                                                          max(float(np.max(np.abs(yc))), 1e-300))}
    del tl
    t_tr, Tt = T.once(lambda: sp.transpose(A))
    out["transpose_csr_ms"] = r4(t_tr)
    del Tt, E
    row = gen.csr_to_coo_rows(rp)
    p = np.random.default_rng(3).permutation(nnz)
    C = sp.coo_from_host(N, row[p], col[p], val[p], ctx)
    del row, p
    t_cc, B = T.once(lambda: sp.coo_csr(C))
    out["coo_csr_ms"] = r4(t_cc)
    out["coo_csr_matches_csr"] = bool(np.array_equal(B.col.numpy(), col) and np.array_equal(B.row.numpy(), rp))
    log("band:", out)
    return out


# ------------------------------------------------------------------ convert
def sec_convert(args, sp, ctx, T, host):
    rp, col, val, x = host
    N = len(rp) - 1
    nnz = len(col)
    from sciwasm.sparse_amd import gen
    row = gen.csr_to_coo_rows(rp)
    p = np.random.default_rng(4).permutation(nnz)
    C = sp.coo_from_host(N, row[p], col[p], val[p], ctx)
    del p
    t_cc, B = T.once(lambda: sp.coo_csr(C))
    ok = bool(np.array_equal(B.row.numpy(), rp))
    del B
    A = sp.csr_from_host(N, rp, col, val, ctx)
    E = sp.csr_ell(A)
    out = {"coo_csr_ms": r4(t_cc), "coo_csr_entries_per_s": r4(nnz / (t_cc * 1e-3)), "row_ptr_ok": ok,
           "csr_ell": "refused by the reference guard" if E is None else "accepted"}
    # COO SpMV of the same matrix (16 B per entry + x + y): the shuffled COO
    # reduces with device atomics, the row-sorted COO with the segmented
    # reduction of its plan; both checked on the first N/8 rows against
    # the CSR bound
    R = N // 8
    pr = val[:rp[R]] * x[col[:rp[R]]]
    lens = np.diff(rp[:R + 1])
    yref = np.add.reduceat(pr, rp[:R])
    sab = np.add.reduceat(np.abs(pr), rp[:R])
    yref[lens == 0] = 0
    sab[lens == 0] = 0
    u = np.finfo(np.float64).eps / 2
    bnd = 2 * (lens + 1) * u / (1 - (lens + 1) * u) * sab + 1e-300
    xv, yv = sp.vec_from_host(x, ctx=ctx), sp.create_vec(N, 0.0, np.float64, ctx=ctx)
    b_coo = 16 * nnz + 16 * N
    Cs = sp.coo_from_host(N, row, col, val, ctx)
    t1 = time.time()
    tl = sp.CsrTiles.from_coo(Cs, 16384, 8192)
    tiles_build_s = time.time() - t1
    t1 = time.time()
    tu = sp.CsrTiles.from_coo(C, 16384, 8192)
    tiles_build_u = time.time() - t1
    for name, pl in (("coo_shuffled_spmv", sp.CooPlan(C)), ("coo_sorted_spmv", sp.CooPlan(Cs)),
                     ("coo_sorted_tiles_spmv", tl), ("coo_shuffled_tiles_spmv", tu)):
        ms = T(lambda: pl.run(xv.vec, yv.vec, beta=0.0), 10, 2)
        y = yv.array()[:R]
        out[name] = {"ms": r4(ms), "gbs": r4(b_coo / (ms * 1e-3) / 1e9),
                     "pct_hbm": r4(100 * b_coo / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS), "plan": pl.info(),
                     "within_csr_bound_rows": R, "within_csr_bound": bool(np.all(np.abs(y - yref) <= bnd))}
        del pl
    out["coo_sorted_tiles_spmv"]["plan_build_s"] = r4(tiles_build_s)
    out["coo_shuffled_tiles_spmv"]["plan_build_s"] = r4(tiles_build_u)
    del C, Cs, tl, tu, row
    log("convert:", out)
    return out


# ------------------------------------------------------------------ ingest
def sec_ingest(args, sp, ctx, T):
    """MatrixMarket text already in HBM -> sorted COO on the device.  A
    symmetric file (lower triangle stored, '%.16e' values, the layout of
    SuiteSparse exports); checked at full size against numpy's own expansion
    and sort of the generator's arrays."""
    from sciwasm.sparse_amd import gen
    rng = np.random.default_rng(args.seed + 9)
    n, nent = args.ingest_rows, args.ingest_entries
    r = rng.integers(0, n, int(nent * 1.05))
    c = rng.integers(0, n, len(r))
    r, c = np.maximum(r, c), np.minimum(r, c)
    _, first = np.unique(r * n + c, return_index=True)
    keep = rng.permutation(first)[:nent]
    r, c = r[keep], c[keep]
    del first, keep
    pool = gen.value_pool(4096, seed=args.seed, style="e16")
    vi = rng.integers(0, len(pool), len(r))
    t0 = time.time()
    text = gen.mm_text(n, n, r + 1, c + 1, pool, vi, symmetry="symmetric")
    log("ingest: %d entries, %.1f MB text generated in %.1fs" % (len(r), len(text) / 1e6, time.time() - t0))
    host = np.frombuffer(text, np.uint8)
    t0 = time.perf_counter()
    d = sp.DeviceArray.from_host(host, ctx=ctx)
    ctx.sync()
    h2d_s = time.perf_counter() - t0
    parse_ms, coo_ms = [], []
    info = A = None
    for _ in range(3):
        if info is not None:
            info.free()
        A = None
        tp, info = T.once(lambda: sp.read_matrix_MM_files(d, ctx))
        tc, A = T.once(lambda: sp.create_COO_from_MM(info, np.float64))
        parse_ms.append(tp)
        coo_ms.append(tc)
    lines, exact = info.stats()
    # full-size check: the expansion and (row, col) order computed by numpy
    pv = np.array([float(t) for t in pool])
    off = r != c
    er = np.concatenate([r, c[off]])
    ec = np.concatenate([c, r[off]])
    ev = np.concatenate([pv[vi], pv[vi][off]])
    order = np.argsort(er.astype(np.int64) * n + ec, kind="stable")
    ok = (A.nnz == len(er) and np.array_equal(A.row.numpy(), er[order]) and
          np.array_equal(A.col.numpy(), ec[order]) and
          np.array_equal(A.val.numpy().view(np.uint64), ev[order].view(np.uint64)))
    tp, tc = min(parse_ms), min(coo_ms)
    out = {"workload": "symmetric MatrixMarket text, N=%d, %d stored entries, '%%.16e' values" % (n, len(r)),
           "text_bytes": len(text), "entries": len(r), "coo_nnz": int(A.nnz), "exact_values": int(exact),
           "parse_ms": r4(tp), "coo_ms": r4(tc), "total_ms": r4(tp + tc),
           "parse_text_gbs": r4(len(text) / (tp * 1e-3) / 1e9),
           "entries_per_s": r4(len(r) / ((tp + tc) * 1e-3)),
           "h2d_s": r4(h2d_s), "text_gbs_with_h2d": r4(len(text) / (h2d_s + (tp + tc) * 1e-3) / 1e9),
           "check_full_size": bool(ok)}
    if not args.no_cpu:
        import oracle
        k = min(len(r), 1000000)
        sample = gen.mm_text(n, n, r[:k] + 1, c[:k] + 1, pool, vi[:k], symmetry="symmetric")
        t0 = time.perf_counter()
        oi = oracle.mm_coo(sample, np.float64)[0]
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"entries_per_s": r4(k / dt), "cores": 1, "kind": "port",
                               "sample": "first %d entries of the same file (%d-entry COO)" % (k, oi["nnz"])}
    info.free()
    log("ingest:", out)
    return out


# ------------------------------------------------------------------ element-wise
def sec_ewise(args, sp, ctx, T, host):
    """Element-wise ops on the headline matrix's values (206M FP64 values):
    out of place and in place (each value read and written once, 16 B), and
    eliminate_zeros on a CSR copy with 10% of its values zeroed."""
    from sciwasm.sparse_amd._lib import call
    rp, col, val, _ = host
    n = len(val)
    a = sp.DeviceArray.from_host(val, ctx=ctx)
    b = sp.DeviceArray(n, np.float64, ctx)
    ops = sp.ewise.OPS
    out = {"values": n}
    for name, src_dst in (("sqrt", (a, b)), ("neg_in_place", (a, a)), ("expm1", (a, b)), ("sin", (a, b))):
        op = ops.index(name.split("_")[0])
        ms = T(lambda: call("sable_ew_values", ctx.handle, 1, op, 0.0, n, src_dst[0].ptr, src_dst[1].ptr),
               max(3, args.steps // 4), 1)
        out[name] = {"ms": r4(ms), "gbs": r4(16.0 * n / (ms * 1e-3) / 1e9),
                     "pct_hbm": r4(100 * 16.0 * n / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS)}
    del a, b
    z = val.copy()
    z[np.random.default_rng(7).random(n) < 0.1] = 0
    A = sp.csr_from_host(len(rp) - 1, rp, col, z, ctx)
    del z
    t, B = T.once(lambda: sp.ewise._compacted(A))
    # algorithmic bytes: read col+val (12 B) and write the kept 12 B, row_ptr twice
    byt = 12.0 * n + 12.0 * B.nnz + 8.0 * (A.N + 1)
    out["eliminate_zeros_csr"] = {"ms": r4(t), "kept": int(B.nnz), "gbs": r4(byt / (t * 1e-3) / 1e9)}
    log("ewise:", out)
    return out


# ------------------------------------------------------------------ reordering
def sec_reorder(args, sp, ctx, T, host):
    """sort_rows_by_nnz on the headline matrix; reorder_NN on a 2D Laplacian
    whose rows were scattered, with the CSR product timed before and after
    (the locality the ordering gives back to the x gathers)."""
    from sciwasm.sparse_amd import gen
    rp, col, val, _ = host
    N = len(rp) - 1
    A = sp.csr_from_host(N, rp, col, val, ctx)
    t_sort, S = T.once(lambda: sp.sort_rows_by_nnz(A))
    out = {"sort_rows_by_nnz": {"N": N, "nnz": len(val), "ms": r4(t_sort),
                                "gbs": r4((24.0 * len(val) + 16.0 * N) / (t_sort * 1e-3) / 1e9),
                                "bins": [S.num_zero_rows, S.num_one_rows, S.num_two_rows, S.num_three_rows]}}
    del A, S
    n = args.reorder_grid
    lrp, lcol, lval = gen.laplacian_2d(n)
    M = len(lrp) - 1
    p = np.random.default_rng(3).permutation(M)
    lens = np.diff(lrp)[p]
    srp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    starts = lrp[p]
    idx = (np.repeat(starts - np.concatenate([[0], np.cumsum(lens)[:-1]]), lens) + np.arange(srp[-1])).astype(np.int64)
    scol, sval = lcol[idx].astype(np.int32), lval[idx]
    L = sp.csr_from_host(M, srp, scol, sval, ctx)
    x = sp.vec_from_host(np.random.default_rng(4).standard_normal(M), ctx=ctx)
    y = sp.create_vec(M, 0.0, np.float64, ctx)

    def spmv_ms(Mx):
        pl = Mx.plan()
        return T(lambda: sp.spmv(Mx, x, y), args.steps, 2), pl
    t_scat, _ = spmv_ms(L)
    t_s2, S = T.once(lambda: sp.sort_rows_by_nnz(L))
    t_nn, R = T.once(lambda: sp.reorder_NN(S, 16))
    t_re, _ = spmv_ms(R)
    byt = 4.0 * (M + 1) + 12.0 * len(sval) + 16.0 * M
    out["reorder_NN"] = {"matrix": "2D Laplacian %dx%d, rows scattered" % (n, n), "N": M, "nnz": len(sval),
                         "sort_ms": r4(t_s2), "reorder_nn_ms": r4(t_nn),
                         "spmv_scattered_ms": r4(t_scat), "spmv_reordered_ms": r4(t_re),
                         "spmv_scattered_gbs": r4(byt / (t_scat * 1e-3) / 1e9),
                         "spmv_reordered_gbs": r4(byt / (t_re * 1e-3) / 1e9)}
    log("reorder:", out)
    return out


# ------------------------------------------------------------------ spts
def sec_spts(args, sp, ctx, T):
    """Lower solves on the 3D Laplacian 128^3 and the random 1M-row lower
    matrix: the row-block schedule (default block size and a sweep), the
    level graph and the tuned choice between them beside it (unless
    --no-spts-graph), sync-free, combined and FP32 + FP64 refinement; residuals against the long-double oracle, and
    the row-block level solve checked bit for bit against the oracle's
    reference-order level solve."""
    import oracle
    from sciwasm.sparse_amd import gen
    out = {}
    for name, make in (("lap3d_128", lambda: gen.laplacian_3d_lower(128)),
                       ("random_lower_1M", lambda: gen.random_lower(1 << 20, 32, seed=7))):
        rp, col, val = make()
        N, nnz = len(rp) - 1, len(col)
        b = np.random.default_rng(8).uniform(-1, 1, N)
        L = sp.csr_from_host(N, rp, col, val, ctx)
        t0 = time.time()
        an = sp.SptsAnalysis(L)
        t_an = time.time() - t0
        bv = sp.vec_from_host(b, ctx=ctx)
        xv = sp.create_vec(N, 0.0, np.float64, ctx=ctx)
        flops = 2 * nnz - N
        bytes_solve = 4 * (N + 1) + 12 * nnz + 16 * N
        res = {"N": N, "nnz": nnz, "nlevels": an.nlevels, "lanes": an.lanes, "analysis_s": r4(t_an),
               "schedule": an.schedule_info()}

        def timed(key, run, reps=10):
            ms = T(run, reps, 2)
            an.status()
            rr = oracle.rel_residual(N, rp, col, val, xv.array(), b)
            res[key] = {"ms": r4(ms), "gflops": r4(flops / (ms * 1e-3) / 1e9),
                        "gbs": r4(bytes_solve / (ms * 1e-3) / 1e9), "rel_residual": rr}
            return ms

        t0 = time.perf_counter()
        nl, lp, perm, order, _ = oracle.level_sets(N, rp, col)
        nrp, nc_, nv = oracle.level_reorder(N, rp, col, val, perm, order)
        yl = oracle.spts_level(N, nrp, nc_, nv, b[perm])
        want = np.empty_like(yl)
        want[perm] = yl
        res["cpu_level_ms"] = r4((time.perf_counter() - t0) * 1e3)
        best = None
        for B in [0] + [int(v) for v in args.spts_blocks.split(",") if v.strip()]:
            an.set_schedule("blocks", B)
            key = "blocks_%d" % an.schedule_info()["block_rows"]
            ms = timed(key, lambda: an.solve_level(bv, xv))
            res[key]["bit_exact_vs_oracle_level"] = bool(np.array_equal(xv.array(), want))
            res[key]["nsteps"] = an.schedule_info()["nsteps"]
            if best is None or ms < best[1]:
                best = (an.schedule_info()["block_rows"], ms)
        an.set_schedule("blocks", best[0])
        res["blocks_best"] = best[0]
        want_sf = oracle.spts_syncfree(N, rp, col, val, b)
        timed("syncfree_blocks", lambda: an.solve_syncfree(bv, xv))
        res["syncfree_blocks"]["bit_exact_vs_oracle_syncfree"] = bool(np.array_equal(xv.array(), want_sf))
        timed("syncfree_indegree", lambda: an.solve_indegree(bv, xv))
        res["syncfree_indegree"]["bit_exact_vs_oracle_syncfree"] = bool(np.array_equal(xv.array(), want_sf))
        timed("combined", lambda: an.solve_combined(bv, xv))
        res["combined"]["bit_exact_vs_oracle_level"] = bool(np.array_equal(xv.array(), want))
        timed("fp32_refine1", lambda: an.solve_refined(bv, xv, iters=1, inner="level"))
        timed("fp32_refine2", lambda: an.solve_refined(bv, xv, iters=2, inner="level"))
        timed("fp32_refine1_syncfree", lambda: an.solve_refined(bv, xv, iters=1, inner="syncfree"))
        # the FP32 solve and its refinement step in one sweep of the row blocks
        timed("fp32_refine1_fused", lambda: an.solve_refined(bv, xv, iters=1, inner="fused"))
        if not args.no_spts_graph:
            an.set_schedule("graph")
            timed("level_graph", lambda: an.solve_level(bv, xv))
            res["level_graph"]["bit_exact_vs_oracle_level"] = bool(np.array_equal(xv.array(), want))
            timed("syncfree_natural", lambda: an.solve_syncfree(bv, xv))
            res["syncfree_natural"]["bit_exact_vs_oracle_syncfree"] = bool(
                np.array_equal(xv.array(), want_sf))
            lanes0 = an.lanes
            t0 = time.time()
            tuned = an.tune()
            tuned["tune_s"] = r4(time.time() - t0)
            timed("level_tuned", lambda: an.solve_level(bv, xv))
            res["level_tuned"]["bit_exact_vs_oracle_level"] = bool(np.array_equal(xv.array(), want))
            res["level_tuned"]["choice"] = tuned
            timed("fp32_refine1_fused_tuned", lambda: an.solve_refined(bv, xv, iters=1, inner="fused"))
            an.set_lanes(lanes0)
            an.set_schedule("blocks")
        else:
            # the level kernels without a graph: one launch per level
            an.set_schedule("launches")
            timed("level_launches", lambda: an.solve_level(bv, xv))
            res["level_launches"]["bit_exact_vs_oracle_level"] = bool(np.array_equal(xv.array(), want))
            timed("fp32_refine1_fused_launches", lambda: an.solve_refined(bv, xv, iters=1, inner="fused"))
            an.set_schedule("blocks")
        t0 = time.perf_counter()
        oracle.spts_syncfree(N, rp, col, val, b)
        res["cpu_serial_ms"] = r4((time.perf_counter() - t0) * 1e3)
        out[name] = res
        log("spts", name, res)
        del an, L
    return out


# ------------------------------------------------------------------ batch
def sec_batch(args, sp, ctx, T, pool):
    import ctypes
    from sciwasm.sparse_amd import _lib as L
    from sciwasm.sparse_amd.serve import BatchSpmvItem, _upload_items
    from sciwasm.sparse_amd._lib import c_vp, c_i64
    ns, rs, rp, col, val = pool
    count = len(ns)
    nnz = int(rp[-1])
    rows = int(rs[-1])
    out = {"matrices": count, "rows": rows, "nnz": nnz}
    for dtype in (np.float64, np.float32):
        vsz = np.dtype(dtype).itemsize
        code = 1 if dtype == np.float64 else 0
        mats, xs, ys = [], [], []
        t0 = time.time()
        for k in range(count):
            a, b = rs[k], rs[k + 1]
            lrp = (rp[a:b + 1] - rp[a]).astype(np.int32)
            A = sp.csr_from_host(int(ns[k]), lrp, col[rp[a]:rp[b]], val[rp[a]:rp[b]].astype(dtype), ctx)
            A.plan()
            mats.append(A)
            xs.append(sp.DeviceArray(int(ns[k]) * 8, dtype, ctx))
            ys.append(sp.DeviceArray(int(ns[k]) * 8, dtype, ctx))
        for xx in xs:
            xx.fill(0.5)
        rowptrs = [(rp[rs[k]:rs[k + 1] + 1] - rp[rs[k]]) for k in range(count)]
        log("batch %s: %d matrices uploaded in %.1fs" % (np.dtype(dtype).name, count, time.time() - t0))
        plans = {}
        for nrhs in (1, 2, 4, 8):
            items, total = [], 0
            for A, xx, yy in zip(mats, xs, ys):
                blk, nb = c_vp(), c_i64()
                L.call("sable_csr_plan_blocks", A.plan().handle, ctypes.byref(blk), ctypes.byref(nb))
                items.append(BatchSpmvItem(A.row.ptr, A.col.ptr, A.val.ptr, blk.value, xx.ptr, yy.ptr, total, 0.0))
                total += nb.value
            d = _upload_items(items, ctx)
            ms_b = T(lambda: L.call("sable_batch_spmv", ctx.handle, code, nrhs, d.ptr, count, total, 0), 10, 2)
            bytes_ = 4 * (rows + count) + (4 + vsz) * nnz + 2 * vsz * rows * nrhs
            res = {"batched_ms": r4(ms_b), "batched_gbs": r4(bytes_ / (ms_b * 1e-3) / 1e9),
                   "products_per_s": r4(count * nrhs / (ms_b * 1e-3))}
            if nrhs == 1:
                def per_matrix():
                    for A, xx, yy in zip(mats, xs, ys):
                        A.plan().run(xx, yy, beta=0.0)
                ms_p = T(per_matrix, 3, 1)
                res["one_launch_per_matrix_ms"] = r4(ms_p)
                res["speedup"] = r4(ms_p / ms_b)
            tr = min(4096, 8192 // nrhs)
            if tr not in plans:
                t0 = time.time()
                plans.clear()
                plans[tr] = (sp.batch_tile_plans(mats, nrhs=nrhs), time.time() - t0)
            res["tiles"] = batch_tiles(sp, ctx, T, mats, xs, ys, nrhs, plans[tr], bytes_, ms_b)
            res["autotuned"] = batch_autotuned(sp, ctx, T, mats, xs, ys, nrhs, bytes_, res, rowptrs)
            out["spmv_%s_nrhs%d" % (np.dtype(dtype).name, nrhs)] = res
            log("batch spmv %s nrhs=%d" % (np.dtype(dtype).name, nrhs), res)
        del plans, mats, xs, ys
    return out


def batch_autotuned(sp, ctx, T, mats, xs, ys, nrhs, bytes_, res, rowptrs):
    """The same batch under the autotuned batch plan (per row-length class,
    the fastest of row-block and tile schedules measured on 16 of the
    class's matrices), checked against the row-block batch's y on the first
    32 matrices (within the CSR bound: tile classes sum rows in LDS-atomic
    order)."""
    ref = [y.numpy()[:A.N * nrhs].astype(np.float64) for A, y in zip(mats[:32], ys[:32])]
    t0 = time.time()
    bp = sp.BatchPlan(mats, nrhs=nrhs, reps=3, sample=16, row_ptrs=rowptrs)
    tune_s = time.time() - t0
    run = bp.bind(xs, ys)
    ms = T(run.run, 10, 2)
    err = 0.0
    for r, A, y in zip(ref, mats[:32], ys[:32]):
        err = max(err, float(np.max(np.abs(y.numpy()[:A.N * nrhs] - r)) / max(float(np.max(np.abs(r))), 1e-300)))
    info = bp.info()
    fmts = {}
    for v in info.values():
        f = v["format"] if v["format"] == "rowblock" else "tiles%d" % v["tile_rows"]
        fmts[f] = fmts.get(f, 0) + v["matrices"]
    best_fixed = min(res["batched_ms"], res["tiles"]["ms"])
    ib = bp.info_batch()
    return {"ms": r4(ms), "gbs": r4(bytes_ / (ms * 1e-3) / 1e9), "classes": len(info),
            "assignment": ib["assignment"], "assignments_ms": {t["assignment"]: r4(t["ms"]) for t in ib["table"]},
            "per_class_matrices_by_schedule": fmts, "tune_s": r4(tune_s),
            "speedup_vs_fixed_schedules": r4(best_fixed / ms), "max_rel_diff_vs_rowblock_first32": err}


def batch_tiles(sp, ctx, T, mats, xs, ys, nrhs, plans_built, bytes_, ms_rowblock):
    """The same batch through per-matrix column-sorted tile plans (row tiles
    of min(4096, 8192 / nrhs) rows, one panel) in one persistent launch;
    checked against the row-block batch's y on the first 32 matrices."""
    from sciwasm.sparse_amd import _lib as L
    from sciwasm.sparse_amd.serve import batch_tiles_items
    plans, build_s = plans_built
    ref = [y.numpy()[:A.N * nrhs].astype(np.float64) for A, y in zip(mats[:32], ys[:32])]
    d, n, total, maxr, dt, _ = batch_tiles_items(plans, xs, ys, 0.0, nrhs)
    code = 1 if dt == np.float64 else 0
    ms = T(lambda: L.call("sable_batch_spmv_tiles", ctx.handle, code, nrhs, d.ptr, n, total, maxr, 0), 10, 2)
    err = 0.0
    for r, A, y in zip(ref, mats[:32], ys[:32]):
        err = max(err, float(np.max(np.abs(y.numpy()[:A.N * nrhs] - r)) / max(float(np.max(np.abs(r))), 1e-300)))
    out = {"ms": r4(ms), "gbs": r4(bytes_ / (ms * 1e-3) / 1e9), "speedup_vs_rowblock_batch": r4(ms_rowblock / ms),
           "tile_rows": maxr, "plan_build_s": r4(build_s), "max_rel_diff_vs_rowblock_first32": err}
    del d
    return out


def sec_batch_tiles(args, sp, ctx, pool):
    """Only the batched tile product of the batch pool, FP64, one right-hand
    side, three launches (and, with --calib-copy, a 1 GiB device copy first):
    a short program for hardware-counter runs of k_batch_tiles."""
    from sciwasm.sparse_amd import _lib as L
    from sciwasm.sparse_amd.serve import batch_tiles_items
    ns, rs, rp, col, val = pool
    mats, xs, ys = [], [], []
    for k in range(len(ns)):
        a, b = rs[k], rs[k + 1]
        lrp = (rp[a:b + 1] - rp[a]).astype(np.int32)
        mats.append(sp.csr_from_host(int(ns[k]), lrp, col[rp[a]:rp[b]], val[rp[a]:rp[b]], ctx))
        x = sp.DeviceArray(int(ns[k]), np.float64, ctx)
        x.fill(0.5)
        xs.append(x)
        ys.append(sp.DeviceArray(int(ns[k]), np.float64, ctx))
    plans = sp.batch_tile_plans(mats)
    d, n, total, maxr, dt, _ = batch_tiles_items(plans, xs, ys)
    if args.calib_copy:
        a = sp.DeviceArray(1 << 27, np.float64, ctx)
        b = sp.DeviceArray(1 << 27, np.float64, ctx)
        a.fill(1.0)
        L.call("sable_memcpy_d2d", ctx.handle, b.ptr, a.ptr, a.nbytes)
        del a, b
    for _ in range(3):
        L.call("sable_batch_spmv_tiles", ctx.handle, 1, 1, d.ptr, n, total, maxr, 0)
    ctx.sync()
    return {"matrices": len(mats), "units": total, "launches": 3}


def make_lower_pool(pool, count):
    """Diagonally dominant lower triangles of the first `count` matrices."""
    from sciwasm.sparse_amd import lower_dominant
    ns, rs, rp, col, val = pool
    out = []
    for k in range(count):
        a, b = rs[k], rs[k + 1]
        lrp = (rp[a:b + 1] - rp[a]).astype(np.int32)
        out.append((int(ns[k]),) + lower_dominant(lrp, col[rp[a]:rp[b]], val[rp[a]:rp[b]]))
    return out


def sec_batch_spts(args, sp, ctx, T, lows):
    import oracle
    Ls, bs, xs = [], [], []
    for n, lrp, lc, lv in lows:
        Ls.append(sp.csr_from_host(n, lrp, lc, lv, ctx))
        bs.append(sp.DeviceArray.from_host(np.full(n, 0.5), ctx=ctx))
        xs.append(sp.DeviceArray(n, np.float64, ctx))
    rows = sum(lo[0] for lo in lows)
    nnz = sum(len(lo[2]) for lo in lows)
    ms_b = T(lambda: sp.batch_spts(Ls, bs, xs, lanes=4), 5, 1)
    ok_b = all(np.array_equal(x.numpy(), oracle_sf(lo, 0.5)) for x, lo in zip(xs[:16], lows[:16]))
    t0 = time.time()
    ans = [sp.SptsAnalysis(Lm, lanes=4) for Lm in Ls]
    t_an = time.time() - t0
    bvs = [sp.sparse_vec_t(b, Lm.N) for b, Lm in zip(bs, Ls)]
    xvs = [sp.sparse_vec_t(x, Lm.N) for x, Lm in zip(xs, Ls)]

    def per():
        for a, b, x in zip(ans, bvs, xvs):
            a.solve_syncfree(b, x)
    out = {"matrices": len(lows), "rows": rows, "nnz": nnz, "batched_syncfree_ms": r4(ms_b),
           "batched_bit_exact_vs_oracle_first16": ok_b, "analysis_s_all": r4(t_an)}
    # the batched level schedule: one launch per level of the deepest matrix,
    # at 1, 4, 8 and 16 lanes per row (the fastest reported)
    for lanes in (1, 4, 8, 16):
        t0 = time.time()
        plan = sp.BatchLevelPlan(Ls, lanes=lanes)
        build_s = time.time() - t0
        run = plan.bind(bs, xs)
        ms = T(run.run, 5, 1)
        out["level_plan_lanes%d" % lanes] = dict(plan.info(), ms=r4(ms), build_s=r4(build_s))
        if "batched_level_ms" not in out or ms < out["batched_level_ms"]:
            out["batched_level_ms"] = r4(ms)
            out["batched_level_bit_exact_vs_oracle_first16"] = all(
                np.array_equal(x.numpy(), oracle_level(lo, 0.5)) for x, lo in zip(xs[:16], lows[:16]))
        del plan, run
    # the same level schedule in FP32 (values rounded), and FP32 with one FP64
    # refinement step fused into its sweep, at the fastest lane count
    best_lanes = min((1, 4, 8, 16), key=lambda g: out["level_plan_lanes%d" % g]["ms"])
    plan = sp.BatchLevelPlan(Ls, lanes=best_lanes)
    L32 = [sp.csr_from_host(n, lrp, lc, lv.astype(np.float32), ctx) for n, lrp, lc, lv in lows]
    plan32 = sp.BatchLevelPlan(L32, lanes=best_lanes)
    b32 = [sp.DeviceArray.from_host(np.full(n, 0.5, np.float32), ctx=ctx) for n, _, _, _ in lows]
    x32 = [sp.DeviceArray(n, np.float32, ctx) for n, _, _, _ in lows]
    run32 = plan32.bind(b32, x32)
    out["batched_level_fp32_ms"] = r4(T(run32.run, 5, 1))
    runr = plan.bind_refined(bs, xs, vals32=[m.val for m in L32])
    out["batched_level_fp32_refine1_fused_ms"] = r4(T(runr.run, 5, 1))
    out["batched_level_fp32_refine1_fused_max_rel_residual_first16"] = max(
        oracle.rel_residual(n, lrp, lc, lv, x.numpy(), np.full(n, 0.5)) for x, (n, lrp, lc, lv) in
        zip(xs[:16], lows[:16]))
    out["batched_level_precision_lanes"] = best_lanes
    del plan, plan32, run32, runr, L32, b32, x32
    # the per-matrix baseline over every matrix: each analysis's own sync-free
    # solve (row blocks: a sentinel fill and one solve launch per matrix) and
    # the natural-order sync-free kernel (the reference's per-matrix unit)
    for name, sched in (("per_matrix_blocks_ms", "blocks"), ("per_matrix_natural_ms", "graph")):
        for a in ans:
            a.set_schedule(sched)
        out[name] = r4(T(per, 3, 1))
    best = min(out["per_matrix_blocks_ms"], out["per_matrix_natural_ms"])
    out["speedup_vs_best_per_matrix"] = r4(best / min(ms_b, out["batched_level_ms"]))
    log("batch spts:", out)
    return out


def oracle_level(lo, bval):
    import oracle
    n, lrp, lc, lv = lo
    b = np.full(n, bval)
    nl, lp, perm, order, _ = oracle.level_sets(n, lrp, lc)
    nrp, nc_, nv = oracle.level_reorder(n, lrp, lc, lv, perm, order)
    yl = oracle.spts_level(n, nrp, nc_, nv, b[perm])
    x = np.empty_like(yl)
    x[perm] = yl
    return x


def oracle_sf(lo, bval):
    import oracle
    n, lrp, lc, lv = lo
    return oracle.spts_syncfree(n, lrp, lc, lv, np.full(n, bval))


# ------------------------------------------------------------------ serve
def serve_accuracy(S, pool, lows, nsmall, count=16):
    """Served results against the FP64 oracle, on `count` SpMV and `count`
    solve requests submitted together: the largest SpMV error relative to
    the row's sum_j |a_ij x_j| (the scale of the CSR bound), and the largest
    solve relative residual ||b - L x|| / (||L|| ||x|| + ||b||)."""
    import oracle
    ns, rs, rp, col, val = pool
    rng = np.random.default_rng(77)
    ids = []
    for k in range(count):
        a, b = rs[k], rs[k + 1]
        n = int(ns[k])
        x = rng.uniform(-1, 1, n)
        ids.append(("spmv", k, x, S.submit(k, "spmv", x)))
        bb = rng.uniform(-1, 1, lows[k][0])
        ids.append(("spts", k, bb, S.submit(nsmall + k, "spts", bb)))
    err, res = 0.0, 0.0
    for op, k, v, rid in ids:
        out, _ = S.wait(rid)
        if op == "spmv":
            a, b = rs[k], rs[k + 1]
            lrp = (rp[a:b + 1] - rp[a]).astype(np.int32)
            y64, ab = oracle.spmv_csr_fp64(int(ns[k]), lrp, col[rp[a]:rp[b]], val[rp[a]:rp[b]], v)
            err = max(err, float(np.max(np.abs(out - y64) / np.maximum(ab, 1e-300))))
        else:
            n, lrp, lc, lv = lows[k]
            res = max(res, oracle.rel_residual(n, lrp, lc, lv, out, v))
    return {"requests": 2 * count, "spmv_max_err_over_abs_row_sum": err, "spts_max_rel_residual": res}


def sec_serve(args, sp, ctx, pool, lows, big):
    out = {}
    rates = [float(r) for r in args.serve_rates.split(",")]
    ns, rs, rp, col, val = pool
    nsmall = min(512, len(ns))
    modes = [m for m in args.serve_modes.split(",") if m]
    for mode in modes:
        # "<mode>+tiles": the batched tile schedule for the pool's SpMV
        # (rows summed in LDS-atomic order); refinement: one FP64 residual +
        # update step after the FP32 solve
        # "<mode>+sf": the batch's solves in one sync-free launch instead of
        # one launch per level
        base, *opts = mode.split("+")
        sweep = "syncfree" if "sf" in opts else "levels"
        S = sp.Server(mode=base, arena_bytes=args.serve_arena_mb << 20, max_batch=args.serve_max_batch, slots=3,
                      refine_iters=args.refine_iters, spts_lanes=args.serve_lanes, group_rhs=args.group_rhs,
                      batch_tiles="tiles" in opts, spts_sweep=sweep, ctx=ctx)
        t0 = time.time()
        for k in range(nsmall):
            a, b = rs[k], rs[k + 1]
            S.add_matrix((rp[a:b + 1] - rp[a]).astype(np.int32), col[rp[a]:rp[b]], val[rp[a]:rp[b]])
        for n, lrp, lc, lv in lows[:nsmall]:
            S.add_matrix(lrp, lc, lv, lower=True)
        if big is not None:
            for _ in range(args.serve_large):
                S.add_matrix(*big)
        out["pool"] = {"general": nsmall, "lower": min(nsmall, len(lows)),
                       "large": args.serve_large if big is not None else 0,
                       "hbm_used_gb": r4(S.info()["hbm_used"] / 1e9)}
        log("serve %s: pool registered in %.1fs" % (mode, time.time() - t0), out["pool"])
        out[mode + "_accuracy"] = serve_accuracy(S, pool, lows, nsmall)
        log("serve %s accuracy:" % mode, out[mode + "_accuracy"])
        res = []
        saturated = 0
        for rate in rates:
            st = S.run_load(rate=rate, duration_s=args.serve_seconds, spts_fraction=0.3, seed=int(rate))
            row = {k: (r4(v) if isinstance(v, float) else v) for k, v in st.items()}
            row["rate"] = rate
            res.append(row)
            log("serve %s rate %.0f:" % (mode, rate), row)
            # past two rates that the server cannot keep up with, higher
            # offered rates only lengthen the queue
            saturated = saturated + 1 if st["solves_per_s"] < 0.8 * rate else 0
            if saturated >= 2:
                break
        out[mode] = res
        S.close()
    best = {m: max(r["solves_per_s"] for r in out[m]) for m in modes}
    out["max_solves_per_s"] = best
    # the service level: the most solves/s served at a p99 latency <= 5 ms
    slo = {m: max([r["solves_per_s"] for r in out[m] if r["p99_us"] <= 5000.0 and r["failed"] == 0] or [0.0])
           for m in modes}
    out["solves_per_s_at_p99_5ms"] = slo
    if "plain" in slo and slo["plain"] > 0:
        out["slo_speedup_vs_plain"] = {m: r4(v / slo["plain"]) for m, v in slo.items()}
    if "mixed" in best and "plain" in best:
        out["speedup_mixed_vs_plain"] = r4(best["mixed"] / max(best["plain"], 1e-9))
    if "mixed+tiles" in best and "plain" in best:
        out["speedup_mixed_tiles_vs_plain"] = r4(best["mixed+tiles"] / max(best["plain"], 1e-9))
    out["config"] = {"max_batch": args.serve_max_batch, "arena_mb_per_slot": args.serve_arena_mb, "slots": 3,
                     "refine_iters": args.refine_iters, "group_rhs": args.group_rhs, "spts_lanes": args.serve_lanes}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=4194304)
    ap.add_argument("--zipf-s", type=float, default=1.5)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--panel-cols", type=str, default="524288",
                    help="comma list of column-panel widths to time (0 = 2 MiB of x); the best is reported")
    ap.add_argument("--no-rowblock", action="store_true", help="skip timing the unblocked row-block kernel")
    ap.add_argument("--tiles", type=str, default="16384x8192,16384x16384",
                    help="comma list of column-sorted tile shapes ROWSxCOLS to time")
    ap.add_argument("--tiles-mode", type=str, default="1",
                    help="tile schedules to time: 0 one launch per panel, 1 one launch over all panels")
    ap.add_argument("--zipf16-tiles", type=str, default="16384x16384,16384x8192",
                    help="tile shapes ROWSxCOLS timed on the zipf(s=1.73) matrix")
    ap.add_argument("--calib-copy", action="store_true",
                    help="first run one 1 GiB device-to-device copy (a plain copy kernel of known size, used to "
                         "scale hardware byte counters)")
    ap.add_argument("--csr-kernel", type=str, default="0,3",
                    help="comma list of CSR block kernels to time (0 basic, 1 staged, 2 persistent, 3 basic with "
                         "non-temporal matrix loads)")
    ap.add_argument("--block-weight", type=str, default="0",
                    help="comma list of row-block weights (nnz + 2*rows) to time with each panel width (0 = 1024)")
    ap.add_argument("--ingest-rows", type=int, default=4194304)
    ap.add_argument("--ingest-entries", type=int, default=8000000)
    ap.add_argument("--reorder-grid", type=int, default=1000)
    ap.add_argument("--sections", type=str,
                    default="c1,addon,band,zipf16,convert,ingest,ewise,reorder,spts,batch,serve",
                    help="comma list of extra sections (empty for the headline only)")
    ap.add_argument("--batch-count", type=int, default=2048)
    ap.add_argument("--serve-rates", type=str,
                    default="2000,10000,30000,50000,60000,70000,80000,85000,90000,95000,100000,110000,120000,240000")
    ap.add_argument("--serve-seconds", type=float, default=1.0)
    ap.add_argument("--serve-large", type=int, default=8)
    ap.add_argument("--serve-modes", type=str, default="plain,batched,mixed,mixed+tiles")
    ap.add_argument("--serve-max-batch", type=int, default=4096)
    ap.add_argument("--serve-lanes", type=int, default=4, help="lanes per row of the served solves")
    ap.add_argument("--serve-arena-mb", type=int, default=2048)
    ap.add_argument("--refine-iters", type=int, default=1)
    ap.add_argument("--spts-blocks", type=str, default="2048,4096,8192,16384",
                    help="comma list of row-block sizes of the triangular solves to time")
    ap.add_argument("--no-spts-graph", action="store_true",
                    help="leave the hipGraph level schedule out of the spts section")
    ap.add_argument("--group-rhs", type=int, default=8)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs the hot path on a single MI355X; "
              "multi-GPU sharding is a design note, not a measured configuration.", file=sys.stderr)
        sys.exit(2)

    import sciwasm.sparse_amd as sp
    ctx = sp.Context.default()
    T = Timer(ctx)
    head, host = headline(args, sp, ctx, T)
    sections = [s for s in args.sections.split(",") if s.strip()]
    extra = {}

    def run(name, fn):
        t0 = time.time()
        try:
            extra[name] = fn()
        except Exception as e:  # a failing section is reported; the headline line is still printed
            extra[name] = {"error": "%s: %s" % (type(e).__name__, e)}
            log(traceback.format_exc())
        log("section %s done in %.1fs" % (name, time.time() - t0))

    if "c1" in sections:
        run("c1", lambda: sec_c1(args, sp, ctx, T))
    if "addon" in sections:
        run("addon", lambda: sec_addon(args, sp, ctx, host))
    if "band" in sections:
        run("band", lambda: sec_band(args, sp, ctx, T))
    if "zipf16" in sections:
        run("zipf16", lambda: sec_zipf16(args, sp, ctx, T))
    if "convert" in sections:
        run("convert", lambda: sec_convert(args, sp, ctx, T, host))
    if "ingest" in sections:
        run("ingest", lambda: sec_ingest(args, sp, ctx, T))
    if "ewise" in sections:
        run("ewise", lambda: sec_ewise(args, sp, ctx, T, host))
    if "reorder" in sections:
        run("reorder", lambda: sec_reorder(args, sp, ctx, T, host))
    if "spts" in sections:
        run("spts", lambda: sec_spts(args, sp, ctx, T))
    pool = lows = None
    if "batch" in sections or "serve" in sections or "batch_tiles" in sections:
        from sciwasm.sparse_amd import gen
        t0 = time.time()
        pool = gen.batch_csr_concat(args.batch_count, seed=args.seed + 4)
        log("pool generated: %d matrices, %d rows, %d nnz in %.1fs" % (len(pool[0]), pool[1][-1], pool[2][-1],
                                                                      time.time() - t0))
        lows = make_lower_pool(pool, args.batch_count) if "batch_tiles" not in sections else None
    if "batch_tiles" in sections:
        run("batch_tiles", lambda: sec_batch_tiles(args, sp, ctx, pool))
    if "batch" in sections:
        run("batch", lambda: sec_batch(args, sp, ctx, T, pool))
        run("batch_spts", lambda: sec_batch_spts(args, sp, ctx, T, lows))
    if "serve" in sections:
        big = (host[0], host[1], host[2]) if args.serve_large > 0 else None
        run("serve", lambda: sec_serve(args, sp, ctx, pool, lows, big))

    gbs, ms = head["gbs"], head["ms"]
    traffic = measured_traffic(head["best"], head["N"])
    out = {
        "metric": "CSR SpMV GB/s (% HBM peak); served solves/s + p50/p99 latency vs plain",
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "fp64 csr spmv y=A.x, zipf(s=%.2f) rows truncated [1,4096], N=%d, nnz=%d, "
                               "uniform columns, N(0,1) values, int32 indices" % (args.zipf_s, head["N"], head["nnz"]),
                   "N": head["N"], "nnz": head["nnz"], "pct_hbm_peak": round(100 * gbs / HBM_PEAK_GBS, 2),
                   "schedule": head["best"], "variants": head["variants"]},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "traffic": (traffic or {}).get("per_product"),
                     "traffic_per_launch": (traffic or {}).get("per_launch"),
                     "kernel": head["variants"][head["best"]].get("kernel", "k_csr_block"),
                     "limiter": (traffic or {}).get("limiter"),
                     "launches_per_step": head["variants"][head["best"]]["launches_per_step"]},
        "cpu_baseline": head["cpu"],
        "check": head["check"],
    }
    if "serve" in extra and "max_solves_per_s" in extra["serve"]:
        sv = extra["serve"]
        out["serving"] = {"max_solves_per_s": sv["max_solves_per_s"],
                          "solves_per_s_at_p99_5ms": sv.get("solves_per_s_at_p99_5ms"),
                          "slo_speedup_vs_plain": sv.get("slo_speedup_vs_plain"),
                          "speedup_mixed_vs_plain": sv.get("speedup_mixed_vs_plain"),
                          "p50_p99_us": {m: [(r["rate"], r["p50_us"], r["p99_us"]) for r in sv[m]]
                                         for m in sv["max_solves_per_s"]}}
    out["extra"] = extra
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
