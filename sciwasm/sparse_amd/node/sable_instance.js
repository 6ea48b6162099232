// sable_instance.js -- the objects the reference's JS layer expects from its
// WebAssembly set-up, backed by the HIP library through sable_addon.node.
//
// seq/src/sparse32.js:1-10 (and sparse64.js) bind three names:
//   memory            Module['wasmMemory']: `.buffer` is the linear memory
//   malloc_instance   instance whose exports are _malloc/_free
//   sparse_instance   instance of sparse_opt_32.wasm / sparse_opt_64.wasm
// par/src/worker32.js:12-16 binds `my_instance` per worker.  `load(bits)`
// returns all of them with the same shapes, so the reference's typed-array
// code keeps working byte for byte on offsets returned by _malloc.
'use strict';

const path = require('path');

let addon = null;

function load(bits) {
  if (bits !== 32 && bits !== 64) throw new Error('bits must be 32 or 64');
  if (addon === null) addon = require(path.join(__dirname, 'sable_addon.node'));
  return {
    memory: addon.wasmMemory,
    malloc_instance: { exports: addon.malloc_exports },
    sparse_instance: { exports: bits === 32 ? addon.sparse_opt_32 : addon.sparse_opt_64 },
    my_instance: { exports: bits === 32 ? addon.par_sparse_opt_32 : addon.par_sparse_opt_64 },
    stats: addon.stats,
  };
}

module.exports = { load };
