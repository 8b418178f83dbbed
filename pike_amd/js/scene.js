'use strict';
// Scene (render/scene.js:8-181): loads a scene JSON
//   { props: [{ uri, transforms: { rotate: [{axis, angle}], scale, translate },
//               skipMaterials: [names], skipMeshes: [names] }],
//     environment: "<RGBE-in-alpha PNG>", atlasRes: 1024 }
// bakes every glTF prop into one triangle list with transforms applied,
// quantizes tangent/bitangent/normal to snorm8 and uv to snorm16 words
// (the reference's truncating quantizers, :97-118), dedupes materials into
// atlas layers (:43-72, :136-141) and reads the environment image.  The
// result has the reference's host representation (attributes[], indices[],
// materials[], texturePacker, env), which scene_desc.js turns into
// pk_scene_desc.  Paths in the JSON are relative to its directory.
// With load(uri, {instancing: true}) props that share a glTF file (and skip
// lists) become placements of one mesh baked in the file's own space, each
// with its prop transform as a 3x4 matrix: scene.meshes [{attributes,
// indices}] and scene.instances [{mesh, material: -1, transform}], the
// scene's own triangle lists empty (pk_bvh_build_instanced builds the two
// levels; the reference bakes every prop, scene.js:159-170).
// Differences from the reference, all fixes: props without transforms use
// the identity; materials are deduplicated by (file, material index) where
// the reference keys them by a per-primitive random id; the bitangent is
// built from the transformed normal and tangent (the reference crosses the
// untransformed ones); environment bytes are the PNG's own (a browser
// canvas may alter them by premultiplying alpha).
const fs = require('fs');
const path = require('path');
const { GLTFLoader } = require('./gltf_loader');
const { TexturePacker } = require('./texture_packer');
const png = require('./png');

// rotateArbitrary (vector.js:237-251): the reference lists this matrix row
// by row and multiplies it as column-major, i.e. it applies the transpose
function rotateArbitrary(v, axis, angle) {
  const [x, y, z] = axis;
  const s = Math.sin(angle), c = Math.cos(angle), oc = 1 - c;
  const m = [oc * x * x + c, oc * x * y - z * s, oc * z * x + y * s, 0,
    oc * x * y + z * s, oc * y * y + c, oc * y * z - x * s, 0,
    oc * z * x - y * s, oc * y * z + x * s, oc * z * z + c, 0,
    0, 0, 0, 1];
  const p = [v[0], v[1], v[2], 1];
  const out = [0, 0, 0];
  for (let i = 0; i < 3; i++) out[i] = m[i] * p[0] + m[4 + i] * p[1] + m[8 + i] * p[2] + m[12 + i] * p[3];
  return out;
}

function normalize(v) {
  const l = Math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  return [v[0] / l, v[1] / l, v[2] / l];
}

function cross(a, b) {
  return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]];
}

// snorm8 x3 in the low three bytes (Int8Array stores ToInt8 of the product)
function quantizeVec3(v) {
  const bytes = new Int8Array(4);
  bytes[0] = v[0] * (v[0] > 0 ? 127 : 128);
  bytes[1] = v[1] * (v[1] > 0 ? 127 : 128);
  bytes[2] = v[2] * (v[2] > 0 ? 127 : 128);
  return new Uint32Array(bytes.buffer);
}

// The prop transform (rotations in order, then uniform scale, then
// translation, as _transform applies them) as a row-major 3x4 matrix.
function instanceTransform(t) {
  const cols = [[1, 0, 0], [0, 1, 0], [0, 0, 1]].map((e) => {
    let r = e;
    for (const rot of t.rotate) r = rotateArbitrary(r, rot.axis, rot.angle);
    return [r[0] * t.scale, r[1] * t.scale, r[2] * t.scale];
  });
  return [cols[0][0], cols[1][0], cols[2][0], t.translate[0],
    cols[0][1], cols[1][1], cols[2][1], t.translate[1],
    cols[0][2], cols[1][2], cols[2][2], t.translate[2]];
}

function quantizeVec2(v) {
  const shorts = new Int16Array(2);
  shorts[0] = v[0] * (v[0] > 0 ? 32767 : 32768);
  shorts[1] = v[1] * (v[1] > 0 ? 32767 : 32768);
  return new Uint32Array(shorts.buffer);
}

class Scene {
  constructor() {
    this.desc = null;
    this.groups = null;
    this.materials = [];
    this.materialIds = {};
    this.texturePacker = null;
  }

  async load(uri, options) {
    const opts = options || {};
    const file = path.resolve(uri);
    const dir = path.dirname(file);
    this.json = JSON.parse(fs.readFileSync(file, 'utf8'));
    const props = this.json.props || [];
    this.groups = props.map((prop) => {
      const g = new GLTFLoader().load(path.join(dir, prop.uri));
      g.transforms = Object.assign({ rotate: [], scale: 1, translate: [0, 0, 0] }, prop.transforms || {});
      g.skipMaterials = new Set(prop.skipMaterials || []);
      g.skipMeshes = new Set(prop.skipMeshes || []);
      return g;
    });
    this.texturePacker = new TexturePacker(this.json.atlasRes === undefined ? 1024 : this.json.atlasRes);
    this.attributes = [];
    this.indices = [];
    this.lights = [];
    if (opts.instancing) this._createInstances(props);
    else this._createBuffers();
    this.env = this.json.environment ? png.decode(fs.readFileSync(path.join(dir, this.json.environment))) : null;
    return this;
  }

  _createSceneMaterial(material) {
    const p = this.texturePacker;
    return {
      diffMap: material.hasBaseColorTexture() ? p.addTexture(material.getBaseColorTexture(), true)
        : p.addColor(material.getBaseColor()),
      metRoughMap: material.hasMetallicRoughnessTexture() ? p.addTexture(material.getMetallicRoughnessTexture())
        : p.addColor(material.getMetallicRoughness()),
      normMap: material.hasNormalTexture() ? p.addTexture(material.getNormalTexture()) : p.addColor([0.5, 0.5, 1]),
      emitMap: material.hasEmissiveTexture() ? p.addTexture(material.getEmissiveTexture(), true)
        : p.addColor(material.getEmissive()),
    };
  }

  _transform(v, t, rotationOnly) {
    let r = v;
    for (const rot of t.rotate) r = rotateArbitrary(r, rot.axis, rot.angle);
    if (rotationOnly) return r;
    return [r[0] * t.scale + t.translate[0], r[1] * t.scale + t.translate[1], r[2] * t.scale + t.translate[2]];
  }

  _createBuffers() {
    let offset = 0;
    for (const group of this.groups) offset = this._bakeGroup(group, group.transforms, this.attributes, this.indices, offset);
  }

  _createInstances(props) {
    const identity = { rotate: [], scale: 1, translate: [0, 0, 0] };
    const meshOf = {};
    this.meshes = [];
    this.instances = [];
    this.groups.forEach((group, gi) => {
      const key = JSON.stringify([props[gi].uri, [...group.skipMaterials].sort(), [...group.skipMeshes].sort()]);
      if (meshOf[key] === undefined) {
        const m = { attributes: [], indices: [] };
        this._bakeGroup(group, identity, m.attributes, m.indices, 0);
        meshOf[key] = m.indices.length ? this.meshes.length : -1;  // nothing left after the skip lists
        if (m.indices.length) this.meshes.push(m);
      }
      if (meshOf[key] >= 0) {
        this.instances.push({ mesh: meshOf[key], material: -1, transform: instanceTransform(group.transforms) });
      }
    });
  }

  // bakes one prop's primitives with transform t into attributes/indices,
  // vertex numbering from `offset`; returns the next offset
  _bakeGroup(group, t, attributes, indices, offset) {
    {
      for (const mesh of group.meshes) {
        if (mesh.desc.name !== undefined && group.skipMeshes.has(mesh.desc.name)) continue;
        for (const primitive of mesh.primitives) {
          const mname = primitive.material.desc.name;
          if (mname !== undefined && group.skipMaterials.has(mname)) continue;
          let matId = this.materialIds[primitive.material.id];
          if (matId === undefined) {
            matId = this.materials.length;
            this.materialIds[primitive.material.id] = matId;
            this.materials.push(this._createSceneMaterial(primitive.material));
          }
          const n = primitive.indexCount();
          for (let i = 0; i + 2 < n; i += 3) {
            indices.push({ i0: offset + primitive.indexAt(i)[0], i1: offset + primitive.indexAt(i + 1)[0],
              i2: offset + primitive.indexAt(i + 2)[0], matId });
          }
          const count = primitive.attributeCount();
          offset += count;
          let lo = null, hi = null;
          for (let i = 0; i < count; i++) {
            const pos = this._transform(primitive.positionAt(i), t, false);
            const normal = normalize(this._transform(primitive.normalAt(i), t, true));
            const st = primitive.tangentAt(i);
            const tangent = normalize(this._transform(st.slice(0, 3), t, true));
            const c = cross(normal, tangent);
            const bitangent = normalize([c[0] * (st[3] || 1), c[1] * (st[3] || 1), c[2] * (st[3] || 1)]);
            attributes.push({ pos, tangent: quantizeVec3(tangent), bitangent: quantizeVec3(bitangent),
              normal: quantizeVec3(normal), uv: quantizeVec2(primitive.uvAt(i)) });
            if (primitive.material.hasEmissive()) {
              lo = lo ? lo.map((x, k) => Math.min(x, pos[k])) : pos.slice();
              hi = hi ? hi.map((x, k) => Math.max(x, pos[k])) : pos.slice();
            }
          }
          if (lo) this.lights.push({ min: lo, max: hi });
        }
      }
    }
    return offset;
  }
}

module.exports = { Scene, quantizeVec3, quantizeVec2, rotateArbitrary, instanceTransform };
