'use strict';
// PostProcessPass (render/passes/postprocess_pass.js, postprocess.wgsl):
// exposure, ACES fitted tonemap, saturation and sRGB encoding.  Headless, so
// it runs on the host over the read-back accumulation and produces RGBA8
// rows (top row first) for a PNG instead of drawing to a canvas.

const ACES_IN = [0.59719, 0.35458, 0.04823, 0.07600, 0.90834, 0.01566, 0.02840, 0.13383, 0.83777];
const ACES_OUT = [1.60475, -0.53108, -0.07367, -0.10208, 1.10813, -0.00605, -0.00327, -0.07276, 1.07602];

// v * M for a WGSL mat3x3 built from 9 scalars in column order
function rowTimesMat(v, m) {
  return [
    v[0] * m[0] + v[1] * m[1] + v[2] * m[2],
    v[0] * m[3] + v[1] * m[4] + v[2] * m[5],
    v[0] * m[6] + v[1] * m[7] + v[2] * m[8],
  ];
}

function rrtOdt(v) {
  const a = v * (v + 0.0245786) - 0.000090537;
  const b = v * (0.983729 * v + 0.4329510) + 0.238081;
  return a / b;
}

function clamp01(x) {
  return Math.min(Math.max(x, 0), 1);
}

function toSrgb(c) {
  return c <= 0.0031308 ? c * 12.92 : 1.055 * Math.pow(c, 1 / 2.4) - 0.055;
}

class PostProcessPass {
  constructor(device, format, context, accumulatePass) {
    this.device = device;
    this.accumulatePass = accumulatePass;
    this.exposure = 1;
    this.saturation = 1;
  }

  static async create(...args) {
    return new PostProcessPass(...args);
  }

  setExposure(v) {
    this.exposure = v;
  }

  setSaturation(v) {
    this.saturation = v;
  }

  generateCommands(commandEncoder) {}

  tonemap(rgba, width, height) {
    const out = new Uint8Array(width * height * 4);
    const s = this.saturation;
    const lum = [0.3086, 0.6094, 0.0820];
    for (let i = 0; i < width * height; i++) {
      let c = [rgba[4 * i] * this.exposure, rgba[4 * i + 1] * this.exposure, rgba[4 * i + 2] * this.exposure];
      c = rowTimesMat(c, ACES_IN).map(rrtOdt);
      c = rowTimesMat(c, ACES_OUT).map(clamp01);
      // saturateColor: mat3x3(red, green, blue) * color
      const r = [lum[0] * (1 - s) + s, lum[0] * (1 - s), lum[0] * (1 - s)];
      const g = [lum[1] * (1 - s), lum[1] * (1 - s) + s, lum[1] * (1 - s)];
      const b = [lum[2] * (1 - s), lum[2] * (1 - s), lum[2] * (1 - s) + s];
      c = [r[0] * c[0] + g[0] * c[1] + b[0] * c[2], r[1] * c[0] + g[1] * c[1] + b[1] * c[2],
        r[2] * c[0] + g[2] * c[1] + b[2] * c[2]];
      for (let k = 0; k < 3; k++) out[4 * i + k] = Math.round(clamp01(toSrgb(c[k])) * 255);
      out[4 * i + 3] = 255;
    }
    return out;
  }

  toRGBA8() {
    const tex = this.accumulatePass.getAccumulateTexture();
    return this.tonemap(tex.read(), tex.width, tex.height);
  }
}

module.exports = { PostProcessPass };
