"""The Node-API drop-in (sciwasm/sparse_amd/node/): the addon that stands in
for the reference's WebAssembly instances.

CPU: node loads the addon without a GPU; it exports the reference's names
(seq/src/sparse_opt_32.wat exports used by seq/src/sparse32.js, and the
parallel tree's exports called from par/src/worker32.js); _malloc/_free and
memcpy work on the linear memory; a kernel call without a GPU fails loudly.

GPU: tests/node/flow.js drives every export the way the reference's JS does
(typed-array views over _malloc offsets); its outputs are compared with the
oracle: bit for bit for the conversions, transposes, DIA/ELL SpMV (serial
and parallel-tree row slices) and the three triangular solves; CSR/COO SpMV
(serial, slices and partial sums) within the summation-order bound.
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from sciwasm.sparse_amd import gen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "sciwasm", "sparse_amd", "node", "sable_addon.node")
FLOW = os.path.join(ROOT, "tests", "node", "flow.js")
NODE = shutil.which("node")

SERIAL = ["spmv_coo", "spmv_coo_wrapper", "spmv_csr", "spmv_csr_wrapper", "spmv_dia", "spmv_dia_wrapper",
          "spmv_ell", "spmv_ell_wrapper", "transpose_csr", "transpose_dia", "count_transpose_cols_ell",
          "transpose_ell", "memcpy", "coo_csr", "num_cols", "csr_ell", "num_diags", "csr_dia", "coo_csc",
          "spts_csc", "spts_csc_wrapper", "mmread", "count_nonzeros_coo", "copy_nonzeros_coo", "copy_nonzeros_csr",
          "count_nonzeros_dia", "copy_nonzeros_dia", "count_nonzeros_ell", "copy_nonzeros_ell",
          "csr_ell_row", "csr_dia_row"]
PARALLEL = ["spmv_coo_wrapper", "spmv_csr_wrapper", "spmv_dia_col_wrapper", "spmv_ell_col_wrapper", "memcpy",
            "sum", "spts_level_csr", "spts_csr_sync_free_wrapper", "spts_csr_level_sync_free_wrapper",
            "spts_csr_opt_level_sync_free_wrapper", "metadata",
            # the SIMD variants par/src/worker32.js calls (:127-399) and the
            # row-major conversions (par/src/sparse32.js:1840-1973)
            "spmv_coo_gs_wrapper", "spmv_csr_gs_wrapper", "spmv_csr_short_rows_wrapper", "spmv_csr_gs_short_wrapper",
            "spmv_csr_unroll2_wrapper", "spmv_csr_unroll3_wrapper", "spmv_csr_unroll4_wrapper",
            "spmv_csr_unroll6_wrapper", "spmv_dia_col_basic_wrapper", "spmv_bdia_col_wrapper", "spmv_dia_wrapper",
            "spmv_ell_col_gs_wrapper", "spmv_bell_col_gs_wrapper", "spmv_ell_wrapper", "spmv_ell_row_gs_wrapper",
            "csr_ell_row", "csr_dia_row"]
# element-wise exports (seq/src/sparse_opt_32.wat, par/src/sparse_opt_32.wat): <op>_<fmt> and self_<op>_<fmt>
EW_OPS = ["expm1", "log1p", "sin", "tan", "pow", "deg2rad", "rad2deg", "sign", "abs", "neg", "sqrt", "ceil",
          "floor", "trunc", "nearest"]
EW = [p + op + "_" + f for f in ("coo", "dia", "ell") for op in EW_OPS for p in ("", "self_")]
SERIAL += EW + ["self_multiply_scalar_coo"]
PARALLEL += EW + [p + "multiply_scalar_" + f for f in ("coo", "dia", "ell") for p in ("", "self_")]

needs_node = pytest.mark.skipif(NODE is None, reason="node is not installed")


@needs_node
def test_addon_loads_and_exports_reference_names():
    assert os.path.exists(ADDON), "addon not built (run __graft_entry__.build())"
    js = """//
const m = require(process.argv[1]).load(32);
const m64 = require(process.argv[1]).load(64);
const out = {
  serial: Object.keys(m.sparse_instance.exports), serial64: Object.keys(m64.sparse_instance.exports),
  parallel: Object.keys(m.my_instance.exports), malloc: Object.keys(m.malloc_instance.exports),
  bytes: m.memory.buffer.byteLength, same: m.memory.buffer === m64.memory.buffer,
};
const M = m.malloc_instance.exports;
const a = M._malloc(100), b = M._malloc(1000), c = M._malloc(8);
out.offsets = [a, b, c];
M._free(b);
out.reuse = M._malloc(900);
new Int32Array(m.memory.buffer, a, 4).set([7, 8, 9, 10]);
m.sparse_instance.exports.memcpy(a, c, 8);
out.copied = Array.from(new Int32Array(m.memory.buffer, c, 2));
try { m.sparse_instance.exports.spmv_csr_wrapper(a, a, a, a, a, 0, 1); out.kernel = 'ran'; }
catch (e) { out.kernel = e.message; }
try { m.sparse_instance.exports.memcpy(0, m.memory.buffer.byteLength - 4, 64); out.oob = 'no error'; }
catch (e) { out.oob = e.message; }
console.log(JSON.stringify(out));
"""
    js = js.replace("process.argv[1]", json.dumps(os.path.join(ROOT, "sciwasm", "sparse_amd", "node",
                                                                  "sable_instance.js")))
    r = subprocess.run([NODE, "-e", js], capture_output=True, text=True, timeout=120, cwd=ROOT,
                       env=dict(os.environ, SABLE_LINEAR_MEMORY_MB="64"))
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert sorted(out["serial"]) == sorted(SERIAL)
    assert sorted(out["serial64"]) == sorted(SERIAL)
    assert sorted(out["parallel"]) == sorted(PARALLEL)
    assert sorted(out["malloc"]) == ["_free", "_malloc"]
    assert out["bytes"] == 64 << 20 and out["same"]
    a, b, c = out["offsets"]
    assert 0 < a < b < c and a % 64 == 0 and b - a >= 100 and c - b >= 1000
    assert out["reuse"] == b  # first fit reuses the freed block
    assert out["copied"] == [7, 8]
    assert out["oob"] == "offset outside the linear memory"
    if out["kernel"] != "ran":  # no GPU here: the error names the HIP failure
        assert "sable_ctx_create" in out["kernel"]


# ------------------------------------------------------------------ GPU flow
def write_inputs(d, dtype, seed=0, workers=3, inner_max=3):
    rng = np.random.default_rng(seed)
    rp, col, _ = gen.laplacian_2d(40)
    N = len(rp) - 1
    nnz = len(col)
    row = gen.csr_to_coo_rows(rp)
    perm = rng.permutation(nnz)
    coo_row, coo_col = row[perm].astype(np.int32), col[perm].astype(np.int32)
    coo_val = rng.uniform(-1, 1, nnz).astype(dtype)
    x = rng.uniform(-1, 1, N).astype(dtype)
    Lrp, Lc, Lv = gen.random_lower(3000, kmax=12, seed=seed + 1, dtype=dtype)
    n = len(Lrp) - 1
    b = rng.uniform(-1, 1, n).astype(dtype)
    nl, lp, lperm, order, _ = oracle.level_sets(n, Lrp, Lc)
    rrp, rc, rv = oracle.level_reorder(n, Lrp, Lc, Lv, lperm, order)
    br = b[lperm]
    # CSC of L: the transpose's CSR, rows ascending within each column, so
    # the diagonal comes first
    ccp, crow, cval = oracle.transpose_csr(n, Lrp, Lc, Lv)
    z_val = coo_val.copy()
    z_val[rng.random(nnz) < 0.3] = 0
    # the same values in CSR order, for copy_nonzeros_csr
    order = np.lexsort((coo_col, coo_row))
    z_csr_val = z_val[order]
    # rows in runs of one, two and three entries after 50 rows of another
    # worker, then longer rows (the short-row variants' layout)
    lens = np.concatenate([rng.integers(4, 9, 50), np.full(100, 1), np.full(150, 2), np.full(150, 3),
                           rng.integers(4, 10, N - 450)])
    sr_row = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    sr_col = np.concatenate([np.sort(rng.choice(N, k, replace=False)) for k in lens]).astype(np.int32)
    sr_val = rng.uniform(-1, 1, len(sr_col)).astype(dtype)
    arrays = dict(coo_row=coo_row, coo_col=coo_col, coo_val=coo_val, x=x, L_row=Lrp, L_col=Lc, L_val=Lv, L_b=b,
                  sr_row=sr_row, sr_col=sr_col, sr_val=sr_val,
                  z_val=z_val, z_csr_val=z_csr_val,
                  Lr_level_ptr=lp, Lr_perm=lperm, Lr_row=rrp, Lr_col=rc, Lr_val=rv, Lr_b=br,
                  Lc_col=ccp, Lc_row=crow, Lc_val=cval)
    for k, v in arrays.items():
        np.ascontiguousarray(v).tofile(os.path.join(d, k + ".bin"))
    with open(os.path.join(d, "mm.mtx"), "wb") as f:
        f.write(mm_text(seed))
    spec = dict(N=N, nnz=nnz, inner_max=inner_max, workers=workers, spts=dict(N=n, nlevels=int(nl)), mm=True,
                short_rows=dict(N=N, one=50, two=150, three=300, four=450))
    with open(os.path.join(d, "inputs.json"), "w") as f:
        json.dump(spec, f)
    return spec, arrays


def mm_text(seed):
    """A symmetric MatrixMarket file with distinct keys and a few explicit zeros."""
    rng = np.random.default_rng(seed + 7)
    n = 300
    r = rng.integers(1, n + 1, 2500)
    c = rng.integers(1, n + 1, 2500)
    r, c = np.maximum(r, c), np.minimum(r, c)
    _, first = np.unique(r * (n + 1) + c, return_index=True)
    first = np.sort(first)
    pool = gen.value_pool(50, seed=seed) + [b"0"]
    return gen.mm_text(n, n, r[first], c[first], pool, rng.integers(0, len(pool), len(first)), symmetry="symmetric")


def row_bound(rp, col, val, x, y0, dtype, extra=0):
    u = np.finfo(dtype).eps / 2
    k = np.diff(rp).astype(np.float64) + 1 + extra
    g = k * u / (1 - k * u)
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    ab = np.abs(val.astype(np.float64) * x.astype(np.float64)[col])
    s = np.bincount(rows, weights=ab, minlength=len(rp) - 1) + np.abs(y0.astype(np.float64))
    return 2 * g * s + np.finfo(dtype).tiny


def read_out(d, name, dtype):
    return np.fromfile(os.path.join(d, "out_" + name + ".bin"), dtype=dtype)


@needs_node
@pytest.mark.gpu
@pytest.mark.parametrize("resident_vectors", ["1", "0"], ids=["resident_xy", "pool_xy"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_node_flow_matches_oracle(tmp_path, dtype, resident_vectors):
    d = str(tmp_path)
    spec, a = write_inputs(d, dtype)
    bits = "32" if dtype == np.float32 else "64"
    print("node flow", bits, "resident vectors", resident_vectors, flush=True)
    r = subprocess.run([NODE, FLOW, d, bits], capture_output=True, text=True, timeout=300, cwd=ROOT,
                       env=dict(os.environ, SABLE_LINEAR_MEMORY_MB="256", SABLE_RESIDENT_VECTORS=resident_vectors))
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    out = json.load(open(os.path.join(d, "outputs.json")))
    N, nnz, im, W = spec["N"], spec["nnz"], spec["inner_max"], spec["workers"]
    i32 = np.int32

    # conversions
    rp, col, val = oracle.coo_csr(N, a["coo_row"], a["coo_col"], a["coo_val"])
    assert np.array_equal(read_out(d, "csr_row", i32), rp)
    assert np.array_equal(read_out(d, "csr_col", i32), col)
    assert np.array_equal(read_out(d, "csr_val", dtype).view(np.uint8), val.view(np.uint8))
    nc, eidx, edata = oracle.csr_ell(N, rp, col, val)
    assert out["ell_nc"] == nc == 5
    assert np.array_equal(read_out(d, "ell_idx", i32), eidx)
    assert np.array_equal(read_out(d, "ell_data", dtype), edata)
    nd, mn = oracle.num_diags(N, rp, col)
    stride = N if dtype == np.float32 else N - mn
    assert out["dia_nd"] == nd == 5 and out["dia_stride"] == stride
    _, _, doff, ddata = oracle.csr_dia(N, rp, col, val, stride=stride)
    assert np.array_equal(read_out(d, "dia_off", i32), doff)
    assert np.array_equal(read_out(d, "dia_data", dtype), ddata)

    # SpMV, y cleared then y += A.x inner_max times.  DIA and ELL keep the
    # reference's per-row order: bit-exact, checksums equal.  CSR and COO sum
    # in another order: within 2 gamma_{k+1} (|y0| + sum |a x|) per product.
    z = np.zeros(N, dtype)
    x = a["x"]
    exact = {
        "ell": oracle.spmv_ell(N, nc, eidx, edata, x, z, im),
        "dia": oracle.spmv_dia(N, stride, doff, ddata, x, z, im),
    }
    exact["par_ell"], exact["par_dia"] = exact["ell"], exact["dia"]
    for name, yref in exact.items():
        y = read_out(d, "y_" + name, dtype)
        assert np.array_equal(y.view(np.uint8), yref.view(np.uint8)), name
        assert out["fletcher_" + name] == oracle.fletcher_sum(yref), name
    y1 = oracle.spmv_csr(N, rp, col, val, x, z, 1)
    yim = oracle.spmv_csr(N, rp, col, val, x, z, im)
    b1 = row_bound(rp, col, val, x, z, dtype)
    bim = b1 * im + row_bound(rp, col, val, x, yim, dtype) * (im - 1)
    close = {"csr": (yim, bim), "csr_again": (yim, bim), "par_csr": (yim, bim), "csr1": (y1, b1),
             "coo": (oracle.spmv_coo(a["coo_row"], a["coo_col"], a["coo_val"], x, z, im), bim),
             # W partial products, each summed into y: W more roundings per row
             "par_coo": (y1, row_bound(rp, col, val, x, z, dtype, extra=W))}
    for name, (yref, bnd) in close.items():
        y = read_out(d, "y_" + name, dtype)
        err = np.abs(y.astype(np.float64) - yref.astype(np.float64))
        assert np.all(err <= bnd), (name, float(err.max()))
    assert out["mirrors_after_spmv"] > 0
    # the unchanged matrix is not uploaded again; a JS write to its values is
    # seen by the next product, which uploads that one array again
    assert out["uploads_again"] == 0
    assert out["write_tracking"] == 1
    assert out["uploads_mod"] == 1
    vmod = read_out(d, "csr_val_mod", dtype)
    assert not np.array_equal(vmod, val)
    ymod = oracle.spmv_csr(N, rp, col, vmod, x, z, 1)
    err = np.abs(read_out(d, "y_csr_mod", dtype).astype(np.float64) - ymod)
    assert np.all(err <= row_bound(rp, col, vmod, x, z, dtype))

    # x and y resident between products: a repeated call uploads neither,
    # host writes to y and then to x are seen (DIA: bit-exact)
    n2 = out["rx_n"]
    roff, rdata = read_out(d, "rx_off", i32), read_out(d, "rx_data", dtype)
    rx, rxm = read_out(d, "rx_x", dtype), read_out(d, "rx_x_mod", dtype)
    assert out["rx_vec_uploads_repeat"] == 0
    assert not np.array_equal(rx, rxm)
    y2 = oracle.spmv_dia(n2, n2, roff, rdata, rx, np.zeros(n2, dtype), 2)
    yw = oracle.spmv_dia(n2, n2, roff, rdata, rx, np.full(n2, 0.5, dtype), 1)
    yx = oracle.spmv_dia(n2, n2, roff, rdata, rxm, yw, 1)
    y3 = oracle.spmv_dia(n2, n2, roff, rdata, rxm, yx, 3)
    for name, yref in (("rx_y2", y2), ("rx_yw", yw), ("rx_yx", yx), ("rx_y3", y3)):
        assert np.array_equal(read_out(d, name, dtype).view(np.uint8), yref.view(np.uint8)), name

    # parallel-tree SIMD variants: CSR/COO within the bound, column-major
    # DIA/ELL bit for bit (each row keeps its stored order)
    for name, (yref, bnd) in {"par_csr_gs": close["csr"], "par_coo_gs": close["coo"]}.items():
        err = np.abs(read_out(d, "y_" + name, dtype).astype(np.float64) - yref.astype(np.float64))
        assert np.all(err <= bnd), name
    for name, base in (("par_dia_basic", "dia"), ("par_bdia", "dia"), ("par_ell_gs", "ell"), ("par_bell_gs", "ell")):
        assert np.array_equal(read_out(d, "y_" + name, dtype).view(np.uint8), exact[base].view(np.uint8)), name
    sr = spec["short_rows"]
    srp, scol, sval = a["sr_row"], a["sr_col"], a["sr_val"]
    ysr = oracle.spmv_csr(N, srp, scol, sval, x, z, im)
    bsr = row_bound(srp, scol, sval, x, z, dtype) * im + row_bound(srp, scol, sval, x, ysr, dtype) * (im - 1)
    for fn in ("spmv_csr_short_rows_wrapper", "spmv_csr_gs_short_wrapper", "spmv_csr_unroll2_wrapper",
               "spmv_csr_unroll3_wrapper", "spmv_csr_unroll4_wrapper", "spmv_csr_unroll6_wrapper"):
        y = read_out(d, "y_" + fn, dtype)
        assert np.all(y[:sr["one"]] == 0), fn  # another worker's rows are not touched
        err = np.abs(y[sr["one"]:].astype(np.float64) - ysr[sr["one"]:])
        assert np.all(err <= bsr[sr["one"]:]), fn
    assert "short-row group sizes" in out["short_rows_mismatch"]
    # row-major storage: conversions and products bit-exact
    _, ri, rd = oracle.csr_ell_row(N, rp, col, val)
    assert np.array_equal(read_out(d, "rell_idx", i32), ri)
    assert np.array_equal(read_out(d, "rell_data", dtype), rd)
    _, roff, rdd = oracle.csr_dia_row(N, rp, col, val)
    assert np.array_equal(read_out(d, "rdia_off", i32), roff)
    assert np.array_equal(read_out(d, "rdia_data", dtype), rdd)
    for name, yref in (("par_ell_row", oracle.spmv_ell_row(N, nc, ri, rd, x, z, inner_max=im)),
                       ("par_ell_row_gs", oracle.spmv_ell_row(N, nc, ri, rd, x, z, gs=True, inner_max=im)),
                       ("par_dia_row", oracle.spmv_dia_row(N, roff, rdd, x, z, inner_max=im))):
        assert np.array_equal(read_out(d, "y_" + name, dtype).view(np.uint8), yref.view(np.uint8)), name

    # transposes
    trp, tc, tv = oracle.transpose_csr(N, rp, col, val)
    assert np.array_equal(read_out(d, "tcsr_row", i32), trp)
    assert np.array_equal(read_out(d, "tcsr_col", i32), tc)
    assert np.array_equal(read_out(d, "tcsr_val", dtype), tv)
    too, tod = oracle.transpose_dia(N, stride, doff, ddata)
    assert np.array_equal(read_out(d, "tdia_off", i32), too)
    assert np.array_equal(read_out(d, "tdia_data", dtype), tod)
    tnc, toi, tod2 = oracle.transpose_ell(N, nc, eidx, edata)
    assert out["tell_nc"] == tnc
    assert np.array_equal(read_out(d, "tell_idx", i32), toi)
    assert np.array_equal(read_out(d, "tell_data", dtype), tod2)

    # triangular solves in the reference's per-row order: bit-exact
    n = spec["spts"]["N"]
    ys = oracle.spts_syncfree(n, a["L_row"], a["L_col"], a["L_val"], a["L_b"])
    assert np.array_equal(read_out(d, "spts_syncfree", dtype), ys)
    bmod = a["L_b"].copy()
    bmod[5] = 0.75
    ysm = oracle.spts_syncfree(n, a["L_row"], a["L_col"], a["L_val"], bmod)
    assert np.array_equal(read_out(d, "spts_syncfree_bmod", dtype), ysm)
    yl = oracle.spts_level(n, a["Lr_row"], a["Lr_col"], a["Lr_val"], a["Lr_b"])
    assert np.array_equal(read_out(d, "spts_level", dtype), yl)
    assert np.array_equal(read_out(d, "spts_combined", dtype), yl)

    # COO -> CSC and the column sweep: bit-exact
    cp, crow, cval = oracle.coo_csc(N, a["coo_row"], a["coo_col"], a["coo_val"])
    assert np.array_equal(read_out(d, "csc_col", i32), cp)
    assert np.array_equal(read_out(d, "csc_row", i32), crow)
    assert np.array_equal(read_out(d, "csc_val", dtype).view(np.uint8), cval.view(np.uint8))
    yc = oracle.spts_csc(n, a["Lc_col"], a["Lc_row"], a["Lc_val"], a["L_b"])
    assert np.array_equal(read_out(d, "spts_csc", dtype).view(np.uint8), yc.view(np.uint8))
    assert np.array_equal(read_out(d, "spts_csc_wrapper", dtype).view(np.uint8), yc.view(np.uint8))

    # element-wise exports: bit-exact except expm1 (JavaScript's Math in
    # the reference: within 1 ulp of the oracle after rounding to the type)
    def ew(name):
        return read_out(d, "ew_" + name, dtype)

    def eq(a, b):
        return np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])

    cv = a["coo_val"]
    with np.errstate(all="ignore"):
        assert eq(ew("sqrt_coo"), oracle.ew_values("sqrt", 0, cv))
        assert eq(ew("deg2rad_coo"), oracle.ew_values("deg2rad", np.pi, cv))
        assert eq(ew("self_sign_coo"), oracle.ew_values("sign", 0, cv))
        assert eq(ew("self_multiply_coo"), oracle.ew_values("multiply", 2.5, cv))
        assert eq(ew("par_ceil_coo"), oracle.ew_values("ceil", 0, cv))
        e1, e2 = ew("expm1_coo"), oracle.ew_values("expm1", 0, cv)
        it = np.int32 if dtype == np.float32 else np.int64
        assert np.all(np.abs(e1.view(it).astype(np.int64) - e2.view(it).astype(np.int64)) <= 2)
        assert eq(ew("neg_dia"), oracle.ew_dia("neg", 0, N, stride, doff, ddata, out=np.full_like(ddata, 7)))
        assert eq(ew("self_sign_dia"), oracle.ew_dia("sign0", 0, N, stride, doff, ddata, out=ddata))
        assert eq(ew("par_abs_dia"), oracle.ew_dia("abs", 0, N, stride, doff, ddata, out=ddata))
        assert eq(ew("nearest_ell"), oracle.ew_values("nearest", 0, edata))
        assert eq(ew("par_mul_ell"), oracle.ew_values("multiply", -2, edata))

    # count_nonzeros_* / copy_nonzeros_*: bit-exact
    zv = a["z_val"]
    assert out["nz_count"] == oracle.count_nonzeros(zv)
    r_, c_, v_ = oracle.copy_nonzeros_coo(a["coo_row"], a["coo_col"], zv)
    assert np.array_equal(read_out(d, "nz_coo_row", i32), r_) and np.array_equal(read_out(d, "nz_coo_col", i32), c_)
    assert np.array_equal(read_out(d, "nz_coo_val", dtype), v_)
    orp, oc, ov = oracle.copy_nonzeros_csr(N, rp, col, a["z_csr_val"])
    assert np.array_equal(read_out(d, "nz_csr_row", i32), orp) and np.array_equal(read_out(d, "nz_csr_col", i32), oc)
    assert np.array_equal(read_out(d, "nz_csr_val", dtype), ov)
    assert out["nz_dia"] == list(oracle.count_nonzeros_dia(N, stride, doff, ddata))
    ooff, odata = oracle.copy_nonzeros_dia(N, stride, doff, ddata)
    assert np.array_equal(read_out(d, "nz_dia_off", i32), ooff)
    assert np.array_equal(read_out(d, "nz_dia_data", dtype), odata)
    assert out["nz_ell"] == list(oracle.count_nonzeros_ell(N, nc, edata))
    cc, od_, oi_ = oracle.copy_nonzeros_ell(N, nc, edata, eidx)
    assert np.array_equal(read_out(d, "nz_ell_data", dtype), od_)
    assert np.array_equal(read_out(d, "nz_ell_idx", i32), oi_)

    # MatrixMarket text -> COO in the linear memory: bit-exact
    text = open(os.path.join(d, "mm.mtx"), "rb").read()
    for pre, lower in (("mm", False), ("mml", True)):
        info, r, c, v = oracle.mm_coo(text, dtype, lower=lower)
        assert np.array_equal(read_out(d, pre + "_row", i32), r)
        assert np.array_equal(read_out(d, pre + "_col", i32), c)
        assert np.array_equal(read_out(d, pre + "_val", dtype).view(np.uint8), v.view(np.uint8))
    assert out["mm"] == dict(N=300, nnz=int(oracle.mm_coo(text, dtype)[0]["nnz"]), field="real",
                             symmetry="symmetric")
    assert out["freed"] == 1
