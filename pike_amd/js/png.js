'use strict';
// PNG decoding (glTF textures, RGBE environment images) and encoding
// (tonemapped output) on Node's zlib, in place of the browser image
// decoding the reference relies on (render/util/utility.js getImage).
// Decodes every standard PNG: gray, gray+alpha, RGB, RGBA and palette
// images at bit depths 1-16, non-interlaced or Adam7, to RGBA8 (16-bit
// samples keep their high byte; tRNS gives transparency).  The RGBA bytes
// are the file's, with no color management or premultiplication.
const zlib = require('zlib');

const SIGNATURE = Buffer.from([137, 80, 78, 71, 13, 10, 26, 10]);
const CHANNELS = { 0: 1, 2: 3, 3: 1, 4: 2, 6: 4 };

let crcTable = null;
function crc32(buf) {
  if (!crcTable) {
    crcTable = new Uint32Array(256);
    for (let n = 0; n < 256; n++) {
      let c = n;
      for (let k = 0; k < 8; k++) c = c & 1 ? 0xedb88320 ^ (c >>> 1) : c >>> 1;
      crcTable[n] = c >>> 0;
    }
  }
  let c = 0xffffffff;
  for (let i = 0; i < buf.length; i++) c = crcTable[(c ^ buf[i]) & 0xff] ^ (c >>> 8);
  return (c ^ 0xffffffff) >>> 0;
}

function paeth(a, b, c) {
  const p = a + b - c;
  const pa = Math.abs(p - a), pb = Math.abs(p - b), pc = Math.abs(p - c);
  if (pa <= pb && pa <= pc) return a;
  return pb <= pc ? b : c;
}

// Reverses the per-row filters of one (sub)image in place; returns the
// unfiltered rows and the offset just past them.
function unfilter(src, offset, rowBytes, rows, bpp) {
  const out = Buffer.alloc(rowBytes * rows);
  for (let y = 0; y < rows; y++) {
    const type = src[offset++];
    const row = y * rowBytes;
    const prev = row - rowBytes;
    for (let x = 0; x < rowBytes; x++) {
      const raw = src[offset + x];
      const a = x >= bpp ? out[row + x - bpp] : 0;
      const b = y > 0 ? out[prev + x] : 0;
      const c = x >= bpp && y > 0 ? out[prev + x - bpp] : 0;
      let v;
      switch (type) {
        case 0: v = raw; break;
        case 1: v = raw + a; break;
        case 2: v = raw + b; break;
        case 3: v = raw + ((a + b) >> 1); break;
        case 4: v = raw + paeth(a, b, c); break;
        default: throw new Error(`png: bad filter type ${type}`);
      }
      out[row + x] = v & 0xff;
    }
    offset += rowBytes;
  }
  return { rows: out, offset };
}

// sample of channel `ch` of pixel x in an unfiltered row, scaled to 8 bits
function sample(rows, rowOff, x, ch, channels, depth, raw) {
  if (depth === 8) return rows[rowOff + x * channels + ch];
  if (depth === 16) return rows[rowOff + (x * channels + ch) * 2];
  const bit = (x * channels + ch) * depth;
  const v = (rows[rowOff + (bit >> 3)] >> (8 - depth - (bit & 7))) & ((1 << depth) - 1);
  return raw ? v : Math.round((v * 255) / ((1 << depth) - 1));
}

function decode(buf) {
  if (buf.length < 8 || !buf.slice(0, 8).equals(SIGNATURE)) throw new Error('png: not a PNG file');
  let pos = 8;
  let hdr = null, palette = null, trns = null;
  const idat = [];
  while (pos + 8 <= buf.length) {
    const len = buf.readUInt32BE(pos);
    const type = buf.toString('latin1', pos + 4, pos + 8);
    const data = buf.slice(pos + 8, pos + 8 + len);
    pos += 12 + len;
    if (type === 'IHDR') {
      hdr = { width: data.readUInt32BE(0), height: data.readUInt32BE(4), depth: data[8], color: data[9],
        interlace: data[12] };
    } else if (type === 'PLTE') {
      palette = data;
    } else if (type === 'tRNS') {
      trns = data;
    } else if (type === 'IDAT') {
      idat.push(data);
    } else if (type === 'IEND') {
      break;
    }
  }
  if (!hdr) throw new Error('png: missing IHDR');
  const channels = CHANNELS[hdr.color];
  if (!channels) throw new Error(`png: unknown color type ${hdr.color}`);
  if (hdr.color === 3 && !palette) throw new Error('png: palette image without PLTE');
  const { width: W, height: H, depth } = hdr;
  const bpp = Math.max(1, (channels * depth) >> 3);
  const data = zlib.inflateSync(Buffer.concat(idat));
  const rgba = new Uint8Array(W * H * 4);
  const put = (rows, rowOff, x, px) => {
    const o = px * 4;
    if (hdr.color === 3) {
      const i = sample(rows, rowOff, x, 0, 1, depth, true);
      rgba[o] = palette[3 * i];
      rgba[o + 1] = palette[3 * i + 1];
      rgba[o + 2] = palette[3 * i + 2];
      rgba[o + 3] = trns && i < trns.length ? trns[i] : 255;
      return;
    }
    const g = (ch) => sample(rows, rowOff, x, ch, channels, depth, false);
    if (channels <= 2) {
      const v = g(0);
      rgba[o] = rgba[o + 1] = rgba[o + 2] = v;
      rgba[o + 3] = channels === 2 ? g(1) : 255;
      if (channels === 1 && trns && trns.length >= 2) {
        const key = trns.readUInt16BE(0);
        if (sample(rows, rowOff, x, 0, 1, depth, true) === (depth === 16 ? key >> 8 : key)) rgba[o + 3] = 0;
      }
    } else {
      rgba[o] = g(0);
      rgba[o + 1] = g(1);
      rgba[o + 2] = g(2);
      rgba[o + 3] = channels === 4 ? g(3) : 255;
      if (channels === 3 && trns && trns.length >= 6) {
        const k = [0, 2, 4].map((i) => trns.readUInt16BE(i));
        const d = depth === 16 ? 8 : 0;
        if (k[0] >> d === rgba[o] && k[1] >> d === rgba[o + 1] && k[2] >> d === rgba[o + 2]) rgba[o + 3] = 0;
      }
    }
  };
  if (hdr.interlace === 0) {
    const rowBytes = Math.ceil((W * channels * depth) / 8);
    const { rows } = unfilter(data, 0, rowBytes, H, bpp);
    for (let y = 0; y < H; y++) for (let x = 0; x < W; x++) put(rows, y * rowBytes, x, y * W + x);
  } else {
    // Adam7: seven reduced images, each filtered on its own
    const passes = [[0, 0, 8, 8], [4, 0, 8, 8], [0, 4, 4, 8], [2, 0, 4, 4], [0, 2, 2, 4], [1, 0, 2, 2], [0, 1, 1, 2]];
    let off = 0;
    for (const [x0, y0, dx, dy] of passes) {
      const pw = Math.ceil((W - x0) / dx), ph = Math.ceil((H - y0) / dy);
      if (pw <= 0 || ph <= 0) continue;
      const rowBytes = Math.ceil((pw * channels * depth) / 8);
      const r = unfilter(data, off, rowBytes, ph, bpp);
      off = r.offset;
      for (let y = 0; y < ph; y++)
        for (let x = 0; x < pw; x++) put(r.rows, y * rowBytes, x, (y0 + y * dy) * W + x0 + x * dx);
    }
  }
  return { width: W, height: H, data: rgba };
}

function chunk(type, data) {
  const head = Buffer.alloc(8);
  head.writeUInt32BE(data.length, 0);
  head.write(type, 4, 'latin1');
  const crc = Buffer.alloc(4);
  crc.writeUInt32BE(crc32(Buffer.concat([head.slice(4), data])), 0);
  return Buffer.concat([head, data, crc]);
}

// RGBA8 (width*height*4 bytes, first row on top) -> PNG file bytes
function encode(width, height, rgba) {
  const ihdr = Buffer.alloc(13);
  ihdr.writeUInt32BE(width, 0);
  ihdr.writeUInt32BE(height, 4);
  ihdr[8] = 8;
  ihdr[9] = 6;
  const raw = Buffer.alloc((width * 4 + 1) * height);
  for (let y = 0; y < height; y++) {
    raw[y * (width * 4 + 1)] = 0;
    Buffer.from(rgba.buffer, rgba.byteOffset + y * width * 4, width * 4).copy(raw, y * (width * 4 + 1) + 1);
  }
  return Buffer.concat([SIGNATURE, chunk('IHDR', ihdr), chunk('IDAT', zlib.deflateSync(raw)),
    chunk('IEND', Buffer.alloc(0))]);
}

module.exports = { decode, encode, crc32 };
