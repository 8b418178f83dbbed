'use strict';
// TexturePacker (render/util/texture_packer.js:56-136): one RGBA8 layer per
// texture or constant color, every layer res x res.  The reference draws
// each image stretched over a WebGL canvas (LINEAR filtering, REPEAT in u,
// CLAMP in v, sRGB textures decoded to linear before filtering) and reads
// the 8-bit framebuffer back; this restates that resampling on the host:
// layer texel (x, y) takes the bilinear sample of the image at its center
// ((x + 0.5) / res, (y + 0.5) / res), in linear space for sRGB ("corrected")
// images, rounded to 8 bits.  Constant colors fill a layer with
// round(255 * c) and alpha 1 (drawColor's clearColor(r, g, b, 1)).
// The layer size is atlasRes, or the largest texture height if that is
// smaller (getResolution, :118-124), 1 when there are only colors.

function srgbToLinear(c) {
  const s = c / 255;
  return s <= 0.04045 ? s / 12.92 : Math.pow((s + 0.055) / 1.055, 2.4);
}
let srgbTable = null;

function unorm8(v) {
  return Math.min(255, Math.max(0, Math.round(v * 255)));
}

class TexturePacker {
  constructor(atlasRes) {
    this.res = atlasRes;
    this.imageSet = [];
    this.imageKeys = {};
    this.maxRes = 1;
  }

  // image: {width, height, data: RGBA8, key}; corrected: sRGB-encoded
  addTexture(image, corrected) {
    const key = `${image.key}|${corrected ? 'srgb' : 'linear'}`;
    if (this.imageKeys[key] === undefined) {
      this.maxRes = Math.max(this.maxRes, image.height);
      this.imageSet.push({ image, corrected: !!corrected });
      this.imageKeys[key] = this.imageSet.length - 1;
    }
    return this.imageKeys[key];
  }

  addColor(color) {
    const key = color.join(' ');
    if (this.imageKeys[key] === undefined) {
      this.imageSet.push(color.slice());
      this.imageKeys[key] = this.imageSet.length - 1;
    }
    return this.imageKeys[key];
  }

  getResolution() {
    if (this.maxRes < this.res) this.res = this.maxRes;
    return [this.res, this.res, this.imageSet.length];
  }

  getPixels() {
    const res = this.getResolution()[0];
    const pixels = new Uint8Array(res * res * 4 * this.imageSet.length);
    this.imageSet.forEach((item, layer) => {
      const out = pixels.subarray(layer * res * res * 4, (layer + 1) * res * res * 4);
      if (Array.isArray(item)) {
        const c = [unorm8(item[0]), unorm8(item[1]), unorm8(item[2]), 255];
        for (let i = 0; i < res * res; i++) out.set(c, 4 * i);
      } else {
        this._resample(item.image, item.corrected, res, out);
      }
    });
    return pixels;
  }

  _resample(img, corrected, res, out) {
    if (corrected && !srgbTable) {
      srgbTable = new Float64Array(256);
      for (let i = 0; i < 256; i++) srgbTable[i] = srgbToLinear(i);
    }
    const W = img.width, H = img.height, d = img.data;
    const texel = (x, y, c) => {
      const xi = ((x % W) + W) % W;          // REPEAT in u
      const yi = Math.min(Math.max(y, 0), H - 1);  // CLAMP in v
      const v = d[(yi * W + xi) * 4 + c];
      return corrected && c < 3 ? srgbTable[v] : v / 255;
    };
    for (let y = 0; y < res; y++) {
      const fy = ((y + 0.5) / res) * H - 0.5;
      const y0 = Math.floor(fy), ty = fy - y0;
      for (let x = 0; x < res; x++) {
        const fx = ((x + 0.5) / res) * W - 0.5;
        const x0 = Math.floor(fx), tx = fx - x0;
        for (let c = 0; c < 4; c++) {
          const top = texel(x0, y0, c) * (1 - tx) + texel(x0 + 1, y0, c) * tx;
          const bot = texel(x0, y0 + 1, c) * (1 - tx) + texel(x0 + 1, y0 + 1, c) * tx;
          out[(y * res + x) * 4 + c] = unorm8(top * (1 - ty) + bot * ty);
        }
      }
    }
  }
}

module.exports = { TexturePacker, srgbToLinear };
