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
  if (m.instances) {
    desc.meshes = m.meshes.map((x) => ({ positions: readArray(dir, x.positions),
      attributes: readArray(dir, x.attributes), triangles: readArray(dir, x.triangles) }));
    desc.instances = m.instances.map((i) => ({ mesh: i.mesh, material: i.material,
      transform: Float64Array.from(i.transform) }));
  }
  return { desc, camera: m.camera, name: m.name };
}

const DTYPES = [[Float64Array, 'f64'], [Float32Array, 'f32'], [Int32Array, 'i32'], [Uint32Array, 'u32'],
  [Uint8Array, 'u8']];

function putArray(dir, name, arr) {
  fs.writeFileSync(path.join(dir, name), Buffer.from(arr.buffer, arr.byteOffset, arr.byteLength));
  const d = DTYPES.find(([T]) => arr instanceof T);
  return { file: name, dtype: d[1] };
}

// Writes a pk_scene_desc-shaped scene (and camera) in the same format.
function saveRawScene(desc, camera, dir, name) {
  fs.mkdirSync(dir, { recursive: true });
  const m = {
    name: name || 'scene',
    positions: putArray(dir, 'positions.bin', desc.positions),
    attributes: putArray(dir, 'attributes.bin', desc.attributes),
    triangles: putArray(dir, 'triangles.bin', desc.triangles),
    materials: putArray(dir, 'materials.bin', desc.materials),
    atlas: { width: desc.atlas.width, height: desc.atlas.height, layers: desc.atlas.layers,
      format: desc.atlas.format, data: putArray(dir, 'atlas.bin', desc.atlas.data) },
    camera,
  };
  if (desc.env) {
    m.env = { width: desc.env.width, height: desc.env.height, format: desc.env.format, sampler: desc.env.sampler,
      data: putArray(dir, 'env.bin', desc.env.data) };
  }
  if (desc.instances && desc.instances.length) {
    m.meshes = desc.meshes.map((x, k) => ({ positions: putArray(dir, `mesh${k}_positions.bin`, x.positions),
      attributes: putArray(dir, `mesh${k}_attributes.bin`, x.attributes),
      triangles: putArray(dir, `mesh${k}_triangles.bin`, x.triangles) }));
    m.instances = desc.instances.map((i) => ({ mesh: i.mesh, material: i.material,
      transform: Array.from(i.transform) }));
  }
  fs.writeFileSync(path.join(dir, 'scene.json'), JSON.stringify(m));
}

module.exports = { loadRawScene, saveRawScene };
