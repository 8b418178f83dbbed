'use strict';
// glTF 2.0 reader for baking scenes into the tracer's flat buffers: the
// role of render/util/gltf_loader.js (GLTFLoader.load :38-64, node
// hierarchy :86-114, accessors :302-343, PBR metallic-roughness materials
// :121-213), written for Node (files instead of fetch, PNG decoding through
// png.js) and with these loader bugs of the reference fixed:
//   * missing roughnessFactor / metallicFactor default to 1 (glTF spec), not 0;
//   * byteStride is read from the accessor's bufferView (the reference looks
//     it up in the accessor table);
//   * computed vertex normals are per vertex (area-weighted face normals of
//     the triangles that use it), not one object shared by all entries;
//   * textures resolve through textures[i].source; normalized integer
//     accessors are scaled to [0, 1] / [-1, 1]; non-indexed primitives use
//     their vertices in order;
//   * a missing tangent is built from an axis not parallel to the normal,
//     and tangent.w (handedness) is taken from the accessor unchanged.
// Supports .gltf (external files or data: URIs) and .glb containers.
// Images must be PNG (JPEG decoding is not part of this build).
const fs = require('fs');
const path = require('path');
const png = require('./png');

const COMPONENT = {
  5120: { size: 1, get: (v, o) => v.getInt8(o), norm: (x) => Math.max(x / 127, -1) },
  5121: { size: 1, get: (v, o) => v.getUint8(o), norm: (x) => x / 255 },
  5122: { size: 2, get: (v, o) => v.getInt16(o, true), norm: (x) => Math.max(x / 32767, -1) },
  5123: { size: 2, get: (v, o) => v.getUint16(o, true), norm: (x) => x / 65535 },
  5125: { size: 4, get: (v, o) => v.getUint32(o, true), norm: (x) => x },
  5126: { size: 4, get: (v, o) => v.getFloat32(o, true), norm: (x) => x },
};
const ELEMENTS = { SCALAR: 1, VEC2: 2, VEC3: 3, VEC4: 4, MAT2: 4, MAT3: 9, MAT4: 16 };
const TRIANGLES = 4;

// 4x4 column-major matrices (glTF convention)
const Mat = {
  identity() {
    return [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1];
  },
  mul(a, b) {
    const r = new Array(16);
    for (let c = 0; c < 4; c++)
      for (let rr = 0; rr < 4; rr++) {
        let s = 0;
        for (let k = 0; k < 4; k++) s += a[4 * k + rr] * b[4 * c + k];
        r[4 * c + rr] = s;
      }
    return r;
  },
  trs(t, q, s) {
    const x = q[0], y = q[1], z = q[2], w = q[3];
    return [
      (1 - 2 * (y * y + z * z)) * s[0], 2 * (x * y + z * w) * s[0], 2 * (x * z - y * w) * s[0], 0,
      2 * (x * y - z * w) * s[1], (1 - 2 * (x * x + z * z)) * s[1], 2 * (y * z + x * w) * s[1], 0,
      2 * (x * z + y * w) * s[2], 2 * (y * z - x * w) * s[2], (1 - 2 * (x * x + y * y)) * s[2], 0,
      t[0], t[1], t[2], 1,
    ];
  },
  point(m, p) {
    return [0, 1, 2].map((r) => m[r] * p[0] + m[4 + r] * p[1] + m[8 + r] * p[2] + m[12 + r]);
  },
  dir(m, d) {
    return [0, 1, 2].map((r) => m[r] * d[0] + m[4 + r] * d[1] + m[8 + r] * d[2]);
  },
  // inverse transpose of the upper 3x3 (normal matrix), row-major 3x3
  normalMatrix(m) {
    const a = m[0], b = m[4], c = m[8], d = m[1], e = m[5], f = m[9], g = m[2], h = m[6], i = m[10];
    const A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    const det = a * A + b * B + c * C;
    const inv = 1 / det;
    // cofactor matrix / det = (M^-1)^T
    return [A * inv, B * inv, C * inv,
      -(b * i - c * h) * inv, (a * i - c * g) * inv, -(a * h - b * g) * inv,
      (b * f - c * e) * inv, -(a * f - c * d) * inv, (a * e - b * d) * inv];
  },
};

const V = {
  sub: (a, b) => [a[0] - b[0], a[1] - b[1], a[2] - b[2]],
  add: (a, b) => [a[0] + b[0], a[1] + b[1], a[2] + b[2]],
  cross: (a, b) => [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]],
  normalize(a) {
    const l = Math.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    return l > 0 ? [a[0] / l, a[1] / l, a[2] / l] : [0, 0, 1];
  },
  mat3(m, v) {
    return [m[0] * v[0] + m[1] * v[1] + m[2] * v[2], m[3] * v[0] + m[4] * v[1] + m[5] * v[2],
      m[6] * v[0] + m[7] * v[1] + m[8] * v[2]];
  },
};

function readUri(uri, dir) {
  const m = /^data:([^;,]*)(;base64)?,(.*)$/s.exec(uri);
  if (m) return m[2] ? Buffer.from(m[3], 'base64') : Buffer.from(decodeURIComponent(m[3]), 'latin1');
  return fs.readFileSync(path.join(dir, decodeURIComponent(uri)));
}

class GLTFMaterial {
  // factors and texture images of one glTF material (index -1: the default)
  constructor(desc, loader, index) {
    this.desc = desc || {};
    this.loader = loader;
    this.id = `${loader.file}#${index}`;
    this.pbr = this.desc.pbrMetallicRoughness || {};
  }

  texture(info) {
    if (!info) return null;
    return this.loader.textureImage(info.index);
  }

  hasBaseColorTexture() { return !!this.pbr.baseColorTexture; }
  getBaseColorTexture() { return this.texture(this.pbr.baseColorTexture); }
  getBaseColor() { return (this.pbr.baseColorFactor || [1, 1, 1, 1]).slice(0, 4); }
  hasMetallicRoughnessTexture() { return !!this.pbr.metallicRoughnessTexture; }
  getMetallicRoughnessTexture() { return this.texture(this.pbr.metallicRoughnessTexture); }
  getMetallicRoughness() {
    const r = this.pbr.roughnessFactor === undefined ? 1 : this.pbr.roughnessFactor;
    const m = this.pbr.metallicFactor === undefined ? 1 : this.pbr.metallicFactor;
    return [0, r, m];
  }
  hasNormalTexture() { return !!this.desc.normalTexture; }
  getNormalTexture() { return this.texture(this.desc.normalTexture); }
  hasEmissive() { return !!this.desc.emissiveFactor; }
  getEmissive() { return (this.desc.emissiveFactor || [0, 0, 0]).slice(0, 3); }
  hasEmissiveTexture() { return !!this.desc.emissiveTexture; }
  getEmissiveTexture() { return this.texture(this.desc.emissiveTexture); }
}

class GLTFPrimitive {
  constructor(desc, mesh) {
    this.desc = desc;
    this.mesh = mesh;
    this.loader = mesh.loader;
    if (desc.mode !== undefined && desc.mode !== TRIANGLES) throw new Error('only TRIANGLES primitives are supported');
    const mi = desc.material === undefined ? -1 : desc.material;
    this.material = this.loader.material(mi);
    this.normalMatrix = Mat.normalMatrix(mesh.matrix);
    this.positions = this.loader.readAccessor(desc.attributes.POSITION);
    this.count = this.positions.length;
    this.indices = desc.indices === undefined
      ? Array.from({ length: this.count }, (_, i) => i)
      : this.loader.readAccessor(desc.indices).map((v) => v[0]);
    this.normals = desc.attributes.NORMAL === undefined ? this._computeNormals()
      : this.loader.readAccessor(desc.attributes.NORMAL);
    this.tangents = desc.attributes.TANGENT === undefined ? null : this.loader.readAccessor(desc.attributes.TANGENT);
    this.uvs = desc.attributes.TEXCOORD_0 === undefined ? null : this.loader.readAccessor(desc.attributes.TEXCOORD_0);
  }

  indexCount() { return this.indices.length; }
  indexAt(i) { return [this.indices[i]]; }
  attributeCount() { return this.count; }
  positionAt(i) { return Mat.point(this.mesh.matrix, this.positions[i]); }
  uvAt(i) { return this.uvs ? this.uvs[i].slice(0, 2) : [0.5, 0.5]; }
  normalAt(i) { return V.normalize(V.mat3(this.normalMatrix, this.normals[i])); }

  // world-space tangent with its handedness in [3]
  tangentAt(i) {
    if (this.tangents) {
      const t = this.tangents[i];
      return V.normalize(Mat.dir(this.mesh.matrix, t)).concat([t[3] === undefined ? 1 : t[3]]);
    }
    const n = this.normalAt(i);
    const axis = Math.abs(n[1]) < 0.999 ? [0, 1, 0] : [1, 0, 0];
    return V.normalize(V.cross(n, axis)).concat([1]);
  }

  _computeNormals() {
    const acc = this.positions.map(() => [0, 0, 0]);
    for (let i = 0; i + 2 < this.indices.length; i += 3) {
      const [a, b, c] = [this.indices[i], this.indices[i + 1], this.indices[i + 2]];
      const p0 = this.positions[a], p1 = this.positions[b], p2 = this.positions[c];
      const n = V.cross(V.sub(p1, p0), V.sub(p2, p0));  // length = 2 x area
      for (const v of [a, b, c]) acc[v] = V.add(acc[v], n);
    }
    return acc.map(V.normalize);
  }
}

class GLTFMesh {
  constructor(desc, matrix, loader) {
    this.desc = desc;
    this.matrix = matrix;
    this.loader = loader;
    this.primitives = desc.primitives.map((p) => new GLTFPrimitive(p, this));
  }

  getPrimitives() { return this.primitives; }
}

class GLTF {
  constructor(meshes, images, asset) {
    this.meshes = meshes;
    this.images = images;
    this.asset = asset;
  }
}

class GLTFLoader {
  constructor() {
    this.manifest = null;
    this.buffers = [];
    this.images = [];
    this.meshes = [];
    this.materials = {};
  }

  load(file) {
    this.file = path.resolve(file);
    const dir = path.dirname(this.file);
    const bytes = fs.readFileSync(this.file);
    let bin = null;
    if (bytes.length >= 12 && bytes.toString('latin1', 0, 4) === 'glTF') {
      if (bytes.readUInt32LE(4) !== 2) throw new Error(`${file}: only glTF 2.0 containers are supported`);
      let pos = 12;
      while (pos + 8 <= bytes.length) {
        const len = bytes.readUInt32LE(pos), type = bytes.readUInt32LE(pos + 4);
        const data = bytes.slice(pos + 8, pos + 8 + len);
        if (type === 0x4e4f534a) this.manifest = JSON.parse(data.toString('utf8'));
        else if (type === 0x004e4942) bin = data;
        pos += 8 + len;
      }
    } else {
      this.manifest = JSON.parse(bytes.toString('utf8'));
    }
    const m = this.manifest;
    if (!m || !m.asset || String(m.asset.version).split('.')[0] !== '2') throw new Error(`${file}: not glTF 2.0`);
    this.buffers = (m.buffers || []).map((b, i) => (b.uri === undefined ? (i === 0 ? bin : null) : readUri(b.uri, dir)));
    this.imageCache = {};
    this.dir = dir;
    const scene = m.scenes ? m.scenes[m.scene === undefined ? 0 : m.scene] : { nodes: [] };
    for (const n of scene.nodes || []) this._visit(n, Mat.identity());
    this.images = Object.keys(this.imageCache).map((k) => this.imageCache[k]);
    return new GLTF(this.meshes, this.images, m.asset);
  }

  _visit(index, parent) {
    const node = this.manifest.nodes[index];
    const local = node.matrix ? node.matrix.slice()
      : Mat.trs(node.translation || [0, 0, 0], node.rotation || [0, 0, 0, 1], node.scale || [1, 1, 1]);
    const m = Mat.mul(parent, local);
    if (node.mesh !== undefined) this.meshes.push(new GLTFMesh(this.manifest.meshes[node.mesh], m, this));
    for (const c of node.children || []) this._visit(c, m);
  }

  material(index) {
    if (!this.materials[index]) {
      const desc = index >= 0 ? this.manifest.materials[index] : null;
      this.materials[index] = new GLTFMaterial(desc, this, index);
    }
    return this.materials[index];
  }

  // decoded RGBA8 image of texture `index` (cached per image)
  textureImage(index) {
    const tex = this.manifest.textures[index];
    const src = tex.source;
    if (!this.imageCache[src]) {
      const img = this.manifest.images[src];
      let bytes;
      if (img.uri !== undefined) {
        bytes = readUri(img.uri, this.dir);
      } else {
        const v = this.manifest.bufferViews[img.bufferView];
        bytes = this.buffers[v.buffer].slice(v.byteOffset || 0, (v.byteOffset || 0) + v.byteLength);
      }
      let decoded;
      try {
        decoded = png.decode(bytes);
      } catch (e) {
        throw new Error(`${this.file}: image ${src} is not a PNG this build can decode (${e.message})`);
      }
      decoded.key = `${this.file}#image${src}`;
      this.imageCache[src] = decoded;
    }
    return this.imageCache[src];
  }

  // accessor -> array of element arrays (numbers), normalized if flagged
  readAccessor(index) {
    const a = this.manifest.accessors[index];
    const comp = COMPONENT[a.componentType];
    if (!comp) throw new Error(`accessor ${index}: unsupported component type ${a.componentType}`);
    const n = ELEMENTS[a.type];
    const out = new Array(a.count);
    if (a.bufferView === undefined) {
      for (let i = 0; i < a.count; i++) out[i] = new Array(n).fill(0);
      return out;
    }
    const bv = this.manifest.bufferViews[a.bufferView];
    const buf = this.buffers[bv.buffer];
    const view = new DataView(buf.buffer, buf.byteOffset + (bv.byteOffset || 0), bv.byteLength);
    const stride = bv.byteStride || comp.size * n;
    const base = a.byteOffset || 0;
    for (let i = 0; i < a.count; i++) {
      const e = new Array(n);
      for (let k = 0; k < n; k++) {
        const x = comp.get(view, base + i * stride + k * comp.size);
        e[k] = a.normalized ? comp.norm(x) : x;
      }
      out[i] = e;
    }
    return out;
  }
}

module.exports = { GLTFLoader, GLTF, Mat };
