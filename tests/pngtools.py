"""Minimal PNG writer/reader for tests (zlib + struct): every color type, bit
depths 1-16, all five row filters and Adam7 interlacing, so the JS decoder's
unfiltering and unpacking are exercised by independently produced files."""
import struct
import zlib

import numpy as np

CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}
ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def _chunk(kind, data):
    c = struct.pack(">I", len(data)) + kind + data
    return c + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def _pack_row(samples, depth):
    """samples: 1-D integer array of one row -> bytes at `depth` bits each."""
    if depth == 8:
        return samples.astype(np.uint8).tobytes()
    if depth == 16:
        return samples.astype(">u2").tobytes()
    out = bytearray((len(samples) * depth + 7) // 8)
    for i, v in enumerate(samples):
        bit = i * depth
        out[bit >> 3] |= int(v) << (8 - depth - (bit & 7))
    return bytes(out)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def _filter_rows(rows, bpp, start_filter):
    out = bytearray()
    prev = bytes(len(rows[0])) if rows else b""
    for y, row in enumerate(rows):
        f = (start_filter + y) % 5
        out.append(f)
        for x in range(len(row)):
            a = row[x - bpp] if x >= bpp else 0
            b = prev[x]
            c = prev[x - bpp] if x >= bpp else 0
            pred = [0, a, b, (a + b) >> 1, _paeth(a, b, c)][f]
            out.append((row[x] - pred) & 0xFF)
        prev = row
    return bytes(out)


def write(samples, color, depth, palette=None, trns=None, interlace=False, start_filter=0):
    """samples: (h, w, channels) integer array in the file's sample units."""
    h, w = samples.shape[:2]
    ch = CHANNELS[color]
    samples = samples.reshape(h, w, ch)
    bpp = max(1, ch * depth // 8)
    ihdr = struct.pack(">IIBBBBB", w, h, depth, color, 0, 0, 1 if interlace else 0)
    if interlace:
        data = b""
        for x0, y0, dx, dy in ADAM7:
            sub = samples[y0::dy, x0::dx]
            if sub.size == 0:
                continue
            rows = [_pack_row(r.reshape(-1), depth) for r in sub]
            data += _filter_rows(rows, bpp, start_filter)
    else:
        rows = [_pack_row(r.reshape(-1), depth) for r in samples]
        data = _filter_rows(rows, bpp, start_filter)
    out = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr)
    if palette is not None:
        out += _chunk(b"PLTE", np.asarray(palette, np.uint8).tobytes())
    if trns is not None:
        out += _chunk(b"tRNS", bytes(trns))
    return out + _chunk(b"IDAT", zlib.compress(data)) + _chunk(b"IEND", b"")


def read_rgba8(buf):
    """Reads 8-bit RGBA, filter-0 PNGs (what the JS encoder writes)."""
    assert buf[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w = 8, b"", 0
    while pos < len(buf):
        n, = struct.unpack(">I", buf[pos:pos + 4])
        kind, data = buf[pos + 4:pos + 8], buf[pos + 8:pos + 8 + n]
        if kind == b"IHDR":
            w, h, depth, color = struct.unpack(">IIBB", data[:10])
            assert depth == 8 and color == 6
        elif kind == b"IDAT":
            idat += data
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = [raw[y * (4 * w + 1):(y + 1) * (4 * w + 1)] for y in range(h)]
    assert all(r[0] == 0 for r in rows)
    return np.frombuffer(b"".join(r[1:] for r in rows), np.uint8).reshape(h, w, 4)
