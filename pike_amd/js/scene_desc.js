'use strict';
// Converts a scene in the reference's host representation
// (render/scene.js:120-181: attributes[] {pos, tangent, bitangent, normal, uv},
// indices[] {i0, i1, i2, matId}, materials[] {diffMap, metRoughMap, normMap,
// emitMap}, texturePacker, env image) into the flat typed arrays of
// pk_scene_desc.  Scenes that already carry a `desc` are passed through.

const PK_TEX_RGBA8 = 0;
const PK_ENV_RGBE8 = 1;
const PK_ENV_SAMPLER_BINS = 0;

function word(v) {
  // scene.js quantizers return a Uint32Array view over the packed bytes
  if (typeof v === 'number') return v >>> 0;
  return v[0] >>> 0;
}

// attributes[] / indices[] of the host representation -> typed arrays
function flatten(attrs, indices) {
  const nv = attrs.length;
  const positions = new Float64Array(nv * 3);
  const attributes = new Uint32Array(nv * 4);
  for (let i = 0; i < nv; i++) {
    const a = attrs[i];
    positions[3 * i] = a.pos[0];
    positions[3 * i + 1] = a.pos[1];
    positions[3 * i + 2] = a.pos[2];
    attributes[4 * i] = word(a.tangent);
    attributes[4 * i + 1] = word(a.bitangent);
    attributes[4 * i + 2] = word(a.normal);
    attributes[4 * i + 3] = word(a.uv);
  }
  const nt = indices.length;
  const triangles = new Int32Array(nt * 4);
  for (let i = 0; i < nt; i++) {
    const t = indices[i];
    triangles[4 * i] = t.i0;
    triangles[4 * i + 1] = t.i1;
    triangles[4 * i + 2] = t.i2;
    triangles[4 * i + 3] = t.matId;
  }
  return { positions, attributes, triangles };
}

function sceneDesc(scene) {
  if (scene.desc) return scene.desc;
  const { positions, attributes, triangles } = flatten(scene.attributes, scene.indices);
  const materials = new Int32Array(scene.materials.length * 4);
  scene.materials.forEach((m, i) => {
    materials.set([m.diffMap, m.metRoughMap, m.normMap, m.emitMap], 4 * i);
  });
  const desc = { positions, attributes, triangles, materials };
  if (scene.instances && scene.instances.length) {
    // instanced scenes (Scene.load(uri, {instancing: true}))
    desc.meshes = scene.meshes.map((m) => flatten(m.attributes, m.indices));
    desc.instances = scene.instances.map((i) => ({ mesh: i.mesh, material: i.material,
      transform: Float64Array.from(i.transform) }));
  }
  if (scene.texturePacker) {
    const res = scene.texturePacker.getResolution();
    desc.atlas = { width: res[0], height: res[1], layers: res[2], format: PK_TEX_RGBA8,
      data: scene.texturePacker.getPixels() };
  }
  if (scene.env) {
    desc.env = { width: scene.env.width, height: scene.env.height, format: PK_ENV_RGBE8,
      sampler: PK_ENV_SAMPLER_BINS, data: scene.env.data };
  }
  scene.desc = desc;
  return desc;
}

module.exports = { sceneDesc, PK_TEX_RGBA8, PK_ENV_RGBE8, PK_ENV_SAMPLER_BINS };
