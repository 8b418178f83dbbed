'use strict';
// Loads the N-API addon (pike_amd/napi/pike_napi.node) that wraps the C ABI
// of include/pk.h.  There is no JavaScript fallback: without the built addon
// (or without a gfx950 device, for device calls) these calls throw.
const path = require('path');

let addon = null;

function load() {
  if (addon) return addon;
  const file = path.join(__dirname, '..', 'napi', 'pike_napi.node');
  try {
    addon = require(file);
  } catch (e) {
    throw new Error(`pike_amd: cannot load ${file} (build it with make -C pike_amd/napi): ${e.message}`);
  }
  return addon;
}

module.exports = { load };
