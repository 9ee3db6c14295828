"""PNG files of a developed picture (darkroom/export/png.js's role), written from the format's specification (RFC 2083 /
ISO 15948): the signature, IHDR (8 or 16 bits a channel, RGBA or RGB), optional iTXt text, an sRGB chunk (the pixels are sRGB-encoded), the rows
each behind a filter byte, deflated into IDAT, and IEND; every chunk closed by the CRC-32 of its type and data.

The row filter is chosen per row as the one whose output has the least sum of absolute values (taking the bytes as
signed), the heuristic the specification suggests; all five filters are written and read back by the tests.
"""
import struct
import zlib

import numpy as np

SIGNATURE = b'\x89PNG\r\n\x1a\n'


def _chunk(kind, data):
    return struct.pack('>I', len(data)) + kind + data + struct.pack('>I', zlib.crc32(kind + data) & 0xFFFFFFFF)


def _paeth(a, b, c):
    p = a.astype(np.int32) + b - c
    pa, pb, pc = np.abs(p - a), np.abs(p - b), np.abs(p - c)
    return np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, c)).astype(np.uint8)


def _filtered(row, prev, bpp, kind):
    """One row's bytes under filter `kind` (0 none, 1 sub, 2 up, 3 average, 4 Paeth), prev the row above (zeros first)."""
    left = np.concatenate([np.zeros(bpp, np.uint8), row[:-bpp]])
    upleft = np.concatenate([np.zeros(bpp, np.uint8), prev[:-bpp]])
    if kind == 0:
        return row
    if kind == 1:
        return row - left
    if kind == 2:
        return row - prev
    if kind == 3:
        return row - ((left.astype(np.uint16) + prev) >> 1).astype(np.uint8)
    return row - _paeth(left, prev, upleft)


def _itxt(keyword, text):
    """An iTXt chunk: the keyword (Latin-1, at most 79 bytes) and NUL, compression flag and method 0, an empty language
    tag and translated keyword (each NUL-ended), the text in UTF-8."""
    return _chunk(b'iTXt', str(keyword)[:79].encode('latin1') + b'\x00' * 5 + str(text).encode('utf8'))


def encode(img, filters='adaptive', level=6, text=()):
    """PNG bytes of an image [h, w, 4] (RGBA) or [h, w, 3] (RGB), row 0 at the top: uint8 for 8 bits a channel, uint16
    for 16 (samples big-endian, as PNG stores them). filters: 'adaptive', or one filter type 0..4 for every row.
    text: (keyword, value) pairs written as iTXt chunks after IHDR."""
    img = np.asarray(img)
    depth = 16 if img.dtype == np.uint16 else 8
    img = np.ascontiguousarray(img, np.uint16 if depth == 16 else np.uint8)
    h, w, ch = img.shape
    if ch not in (3, 4):
        raise ValueError('png: 3 or 4 channels, not %d' % ch)
    rows = img.astype('>u2').view(np.uint8).reshape(h, -1) if depth == 16 else img.reshape(h, -1)
    bpp = ch * depth // 8                                # the filters look back a whole pixel
    raw = bytearray()
    prev = np.zeros(rows.shape[1], np.uint8)
    for y in range(h):
        row = rows[y]
        if filters == 'adaptive':
            best = min(range(5), key=lambda k: int(np.abs(_filtered(row, prev, bpp, k).view(np.int8).astype(np.int32)).sum()))
        else:
            best = int(filters)
        raw.append(best)
        raw += _filtered(row, prev, bpp, best).tobytes()
        prev = row
    ihdr = struct.pack('>IIBBBBB', w, h, depth, 6 if ch == 4 else 2, 0, 0, 0)
    return SIGNATURE + _chunk(b'IHDR', ihdr) + b''.join(_itxt(k, v) for k, v in text) + _chunk(b'sRGB', b'\x00') + \
        _chunk(b'IDAT', zlib.compress(bytes(raw), level)) + _chunk(b'IEND', b'')
