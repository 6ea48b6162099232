// bench_node.js -- times the reference's own call surface through the
// Node-API addon: a CSR matrix in the linear memory (arrays at _malloc
// offsets, filled through typed-array views, as seq/src/sparse32.js does),
// then spmv_csr_wrapper(rowptr, col, val, x, y, N, inner_max) calls, each a
// full y += A.x inner_max times, y brought back to the linear memory
// (seq/tests/main32.js:122-134 times the same call after its warm-up).
// `uploads` counts matrix-array uploads and `vector_uploads` uploads of x
// or y (both stay on the device while the caller leaves them unchanged).
//
// usage: node bench_node.js <dir> <32|64> <warmup> <calls> <inner_max>[,<inner_max>...]
// <dir> holds spec.json ({N, nnz}) and rp.bin, col.bin, val.bin, x.bin
// (raw little-endian); prints one JSON line and writes <dir>/y.bin, the
// result of one further call with y cleared and inner_max = 1.
'use strict';

const fs = require('fs');
const path = require('path');

const dir = process.argv[2];
const bits = parseInt(process.argv[3], 10);
const warmup = parseInt(process.argv[4], 10);
const calls = parseInt(process.argv[5], 10);
const inner = process.argv[6].split(',').map((v) => parseInt(v, 10));
const inst = require(path.join(__dirname, 'sable_instance.js')).load(bits);
const E = inst.sparse_instance.exports;
const M = inst.malloc_instance.exports;
const FArray = bits === 32 ? Float32Array : Float64Array;
const spec = JSON.parse(fs.readFileSync(path.join(dir, 'spec.json'), 'utf8'));
const N = spec.N, nnz = spec.nnz;

// a file into the linear memory, read straight into the typed-array view
function load(name, count, Type) {
  const bytes = count * Type.BYTES_PER_ELEMENT;
  const off = M._malloc(Math.max(bytes, 4));
  if (!off) throw new Error('linear memory too small for ' + name);
  const fd = fs.openSync(path.join(dir, name + '.bin'), 'r');
  const view = new Uint8Array(inst.memory.buffer, off, bytes);
  for (let done = 0; done < bytes;) {
    const n = fs.readSync(fd, view, done, Math.min(bytes - done, 1 << 30), done);
    if (n <= 0) throw new Error('short read of ' + name);
    done += n;
  }
  fs.closeSync(fd);
  return off;
}

const t0 = process.hrtime.bigint();
const rp = load('rp', N + 1, Int32Array);
const col = load('col', nnz, Int32Array);
const val = load('val', nnz, FArray);
const x = load('x', N, FArray);
const y = M._malloc(N * FArray.BYTES_PER_ELEMENT);
const yv = new FArray(inst.memory.buffer, y, N);
const load_s = Number(process.hrtime.bigint() - t0) / 1e9;

const ms = (a, b) => Number(b - a) / 1e6;
const out = { N: N, nnz: nnz, load_s: load_s, runs: {} };
// the first call uploads the matrix and builds its schedule
yv.fill(0);
let a = process.hrtime.bigint();
E.spmv_csr_wrapper(rp, col, val, x, y, N, 1);
out.first_call_ms = ms(a, process.hrtime.bigint());
out.schedule = inst.stats().schedule;
out.schedule_ms = inst.stats().schedule_ms;
for (const im of inner) {
  for (let i = 0; i < warmup; i++) E.spmv_csr_wrapper(rp, col, val, x, y, N, im);
  const up0 = inst.stats().uploads, vu0 = inst.stats().vector_uploads;
  const t = [];
  for (let i = 0; i < calls; i++) {
    a = process.hrtime.bigint();
    E.spmv_csr_wrapper(rp, col, val, x, y, N, im);
    t.push(ms(a, process.hrtime.bigint()));
  }
  t.sort((p, q) => p - q);
  out.runs[im] = { calls: calls, median_ms: t[Math.floor(calls / 2)], min_ms: t[0],
                   mean_ms: t.reduce((p, q) => p + q, 0) / calls, uploads: inst.stats().uploads - up0,
                   vector_uploads: inst.stats().vector_uploads - vu0 };
}
yv.fill(0);
E.spmv_csr_wrapper(rp, col, val, x, y, N, 1);
fs.writeFileSync(path.join(dir, 'y.bin'), Buffer.from(new Uint8Array(inst.memory.buffer, y, N * FArray.BYTES_PER_ELEMENT)));
const st = inst.stats();
out.write_tracking = st.write_tracking;
out.pinned = st.pinned;
console.log(JSON.stringify(out));
