'use strict';
// EnvironmentGenerator (render/env_sampler.js:7-201): importance-sampling
// tables of an RGBE-in-alpha environment image {width, height, data (RGBA8)}
// or a float RGB map.  The tables are built natively (the same code the
// scene upload uses), so what JS inspects is what the kernels sample.
const native = require('./native');
const { sceneDesc } = require('./scene_desc');

class EnvironmentGenerator {
  constructor(img, options) {
    this.img = img;
    const opts = options || {};
    this.format = opts.format === undefined ? 1 : opts.format;
    this.sampler = opts.sampler || 0;
    this._tables = null;
  }

  _build() {
    if (this._tables) return this._tables;
    const desc = {
      positions: new Float64Array(9),
      triangles: new Int32Array([0, 1, 2, 0]),
      env: { width: this.img.width, height: this.img.height, format: this.format, sampler: this.sampler,
        data: this.img.data },
    };
    this._tables = native.load().envTables(desc);
    return this._tables;
  }

  // {h0, h1} pairs (LuminanceBinStruct)
  createHistogramBuffer() {
    return this._build().bins;
  }

  // {x, y} texel coordinates by descending radiance (LuminanceCoordStruct)
  createEnvCoordBuffer() {
    return this._build().coords;
  }

  // per-texel pdf, row major (r32float texture of the reference)
  createPdfTexture() {
    return this._build().pdf;
  }
}

module.exports = { EnvironmentGenerator, sceneDesc };
