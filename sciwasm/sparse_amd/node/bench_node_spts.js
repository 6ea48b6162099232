// bench_node_spts.js -- times the reference's sync-free lower solve through
// the Node-API addon: a lower-triangular CSR (diagonal last in each row) in
// the linear memory, then calls of the parallel tree's
// spts_csr_sync_free_wrapper(id, rowptr, col, val, b, y, N, barrier,
// global_flag, flag, nthreads, inner_max) from worker 0 (the other workers'
// calls return at once; par/src/spts_opt_32.wat:119-260 is the unit it
// replaces).  b stays on the device while unchanged; y comes back every call.
//
// usage: node bench_node_spts.js <dir> <32|64> <warmup> <calls>
// <dir> holds spec.json ({N, nnz}) and rp.bin, col.bin, val.bin, b.bin;
// prints one JSON line and writes <dir>/y.bin (the last call's solution).
'use strict';

const fs = require('fs');
const path = require('path');

const dir = process.argv[2];
const bits = parseInt(process.argv[3], 10);
const warmup = parseInt(process.argv[4], 10);
const calls = parseInt(process.argv[5], 10);
const inst = require(path.join(__dirname, 'sable_instance.js')).load(bits);
const P = inst.my_instance.exports;
const M = inst.malloc_instance.exports;
const FArray = bits === 32 ? Float32Array : Float64Array;
const spec = JSON.parse(fs.readFileSync(path.join(dir, 'spec.json'), 'utf8'));
const N = spec.N, nnz = spec.nnz;

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

const rp = load('rp', N + 1, Int32Array);
const col = load('col', nnz, Int32Array);
const val = load('val', nnz, FArray);
const b = load('b', N, FArray);
const y = M._malloc(N * FArray.BYTES_PER_ELEMENT);
const barrier = M._malloc(4), gflag = M._malloc(4), flag = M._malloc(4 * N);
const solve = () => P.spts_csr_sync_free_wrapper(0, rp, col, val, b, y, N, barrier, gflag, flag, 1, 1);
const ms = (a, c) => Number(c - a) / 1e6;

const out = { N: N, nnz: nnz };
let a = process.hrtime.bigint();
solve();  // mirrors the matrix and b, builds the analysis
out.first_call_ms = ms(a, process.hrtime.bigint());
for (let i = 0; i < warmup; i++) solve();
const up0 = inst.stats().uploads, vu0 = inst.stats().vector_uploads;
const t = [];
for (let i = 0; i < calls; i++) {
  a = process.hrtime.bigint();
  solve();
  t.push(ms(a, process.hrtime.bigint()));
}
t.sort((p, q) => p - q);
out.calls = calls;
out.median_ms = t[Math.floor(calls / 2)];
out.min_ms = t[0];
out.uploads = inst.stats().uploads - up0;
out.vector_uploads = inst.stats().vector_uploads - vu0;
fs.writeFileSync(path.join(dir, 'y.bin'), Buffer.from(new Uint8Array(inst.memory.buffer, y, N * FArray.BYTES_PER_ELEMENT)));
console.log(JSON.stringify(out));
