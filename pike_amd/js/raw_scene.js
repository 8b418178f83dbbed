'use strict';
// Reads a scene stored as a JSON manifest plus raw little-endian arrays
// (the format pike_amd.scenes.SceneData.save writes): lets JavaScript and
// Python drive the same bytes through the same C ABI.
const fs = require('fs');
const path = require('path');

const TYPES = {
  f64: Float64Array, f32: Float32Array, i32: Int32Array, u32: Uint32Array, u8: Uint8Array,
};

function readArray(dir, entry) {
  const buf = fs.readFileSync(path.join(dir, entry.file));
  const T = TYPES[entry.dtype];
  const copy = new Uint8Array(buf.length);
  copy.set(buf);
  return new T(copy.buffer, 0, buf.length / T.BYTES_PER_ELEMENT);
}

function loadRawScene(dir) {
  const m = JSON.parse(fs.readFileSync(path.join(dir, 'scene.json'), 'utf8'));
  const desc = {
    positions: readArray(dir, m.positions),
    attributes: readArray(dir, m.attributes),
    triangles: readArray(dir, m.triangles),
    materials: readArray(dir, m.materials),
    atlas: Object.assign({}, m.atlas, { data: readArray(dir, m.atlas.data) }),
  };
  if (m.env) desc.env = Object.assign({}, m.env, { data: readArray(dir, m.env.data) });
  return { desc, camera: m.camera, name: m.name };
}

module.exports = { loadRawScene };
