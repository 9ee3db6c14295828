"""The developed picture's checker (oracle/develop.py) against what fixes it independently -- the sRGB standard's
values, each tone curve's defining property, the photographic meter's closed form -- and the PNG writers against a
decoder written here from the format's specification."""
import json
import math
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

import develop as D
from shadytracer_amd import png

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def grey(x):
    return np.array([x, x, x], float)


def test_srgb_encoding_values():
    # IEC 61966-2-1: the linear toe meets the power curve at 0.0031308 -> 0.04045; 0.5 -> 0.7353569; 0.18 -> 0.4613561
    assert D.srgb(0.0031308) == pytest.approx(0.04045, abs=2e-6)
    assert D.srgb(0.5) == pytest.approx(0.7353569, abs=1e-6)
    assert D.srgb(0.18) == pytest.approx(0.4613561, abs=1e-6)
    assert D.srgb(1.0) == pytest.approx(1.0) and D.srgb(2.0) == pytest.approx(1.0) and D.srgb(-1.0) == 0.0


def test_tone_curves_defining_properties():
    # Reinhard on luminance: L / (1 + L), the hue kept
    c = np.array([2.0, 1.0, 0.5])
    L = c @ D.LUMA
    out = D.tone_reinhard(c)
    assert out @ D.LUMA == pytest.approx(L / (1 + L)) and np.allclose(out / out[0], c / c[0])
    # PBR Neutral: below the compression start only the toe offset (0.04 above 0.08); continuous at the start; never
    # above 1; a blinding input goes to white
    assert np.allclose(D.tone_neutral(grey(0.3)), grey(0.26))
    assert np.allclose(D.tone_neutral(grey(0.05)), grey(0.05 - (0.05 - 6.25 * 0.05 ** 2)))
    a, b = D.tone_neutral(grey(0.76 + 0.04 - 1e-9)), D.tone_neutral(grey(0.76 + 0.04 + 1e-9))
    assert np.allclose(a, b, atol=1e-7)
    assert D.tone_neutral(np.array([1e6, 10.0, 1.0])).max() <= 1.0
    assert np.ptp(D.tone_neutral(np.array([1e6, 1e5, 1e4]))) < 0.01
    # Hable: white (11.2, after the bias of 2) maps to 1
    assert np.allclose(D.tone_hable(grey(5.6)), 1.0)
    # Hill's ACES fit keeps grey grey (each matrix's rows sum to 1: a transposed matrix would not) and is the rational
    # fit on it: 1.0 -> (1.0245786 - 0.000090537) / (0.983729 + 0.4329510 + 0.238081)
    assert np.allclose(D.tone_aces(grey(1.0)), (1.0245786 - 0.000090537) / (0.983729 + 0.4329510 + 0.238081), atol=1e-4)
    # AgX puts mid grey near the middle of the display range, and grey stays (nearly) grey
    g = D.srgb(D.tone_agx(grey(0.18)))
    assert g.mean() == pytest.approx(0.50, abs=0.01) and np.ptp(g) < 2e-3
    for k, f in D.TONES.items():
        xs = np.stack([grey(x) for x in np.geomspace(1e-3, 1e3, 200)])
        ys = f(xs)[:, 1] if k else xs[:, 1]
        assert np.all(np.diff(ys) >= -1e-12), k      # monotone


def test_meter_on_a_uniform_picture():
    # every pixel at L: the trimmed average is L's bin, the meter EV = log2(L 100 / 12.5) within half a bin (1/6 stop)
    for L in (0.01, 3.0, 800.0, 2.0e5):
        acc = np.zeros((20, 30, 4))
        acc[..., :3], acc[..., 3] = L, 4
        ev, logl, _ = D.ae_update(D.ae_histogram(acc))
        assert ev == pytest.approx(math.log2(L * 8.0), abs=44 / 128 / 2 + 1e-9), L
    # nothing above the floor: no reading
    assert D.ae_update(D.ae_histogram(np.zeros((4, 4, 4))))[0] is None


def test_meter_guards_the_highlights():
    # 92 % of the pixels at 1 cd/m^2 and 8 % at 1e6 (a sun in the frame): the trimmed average meters the room (EV 3),
    # but the 95 % quantile then clips by 18 stops, so the picture is darkened until it sits `headroom` stops above
    # white: EV = log2(1e6) - log2(1.2) - headroom
    acc = np.zeros((20, 20, 4))
    acc[..., :3], acc[..., 3] = 1.0, 1
    acc.reshape(-1, 4)[:32, :3] = 1e6
    ev, logl, logq = D.ae_update(D.ae_histogram(acc), headroom=2.0)
    assert logl == pytest.approx(0.0, abs=0.2) and logq == pytest.approx(math.log2(1e6), abs=0.35)
    assert ev == pytest.approx(math.log2(1e6) - math.log2(1.2) - 2.0, abs=0.35)
    # easing: a quarter of the way from the last exposure
    ev2, _, _ = D.ae_update(D.ae_histogram(acc), headroom=2.0, speed=0.25, state=0.0)
    assert ev2 == pytest.approx(0.25 * ev)


def test_develop_exposes_white():
    # at the metered EV a uniform picture lands at 12.5 / (100 x 1.2) of white, linear (the 18 %-card exposure)
    acc = np.zeros((8, 8, 4))
    acc[..., :3], acc[..., 3] = 800.0, 1
    rgb, v, ev = D.develop(acc, tone=0)
    lin = 800.0 / (1.2 * 2.0 ** ev)
    assert lin == pytest.approx(12.5 / 120.0, rel=0.13)
    assert np.all(rgb == np.floor(D.srgb(lin) * 255 + 0.5))


def decode_png(data):
    """A reader written from the specification: chunks with their CRCs, IHDR, IDAT concatenated and inflated, every
    row unfiltered by its filter type."""
    assert data[:8] == b'\x89PNG\r\n\x1a\n'
    pos, idat, ihdr, kinds = 8, b'', None, []
    while pos < len(data):
        n, = struct.unpack('>I', data[pos:pos + 4])
        kind, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        crc, = struct.unpack('>I', data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(kind + body) & 0xFFFFFFFF, kind
        kinds.append(kind)
        if kind == b'IHDR':
            ihdr = struct.unpack('>IIBBBBB', body)
        elif kind == b'IDAT':
            idat += body
        pos += 12 + n
    w, h, depth, ctype = ihdr[:4]
    assert depth in (8, 16) and ctype in (2, 6) and kinds[-1] == b'IEND'
    # filters work on bytes, looking back a whole pixel: bpp bytes
    ch = (4 if ctype == 6 else 3) * depth // 8
    raw = zlib.decompress(idat)
    out = np.zeros((h, w * ch), np.int32)
    prev = np.zeros(w * ch, np.int32)
    for y in range(h):
        f = raw[y * (w * ch + 1)]
        row = np.frombuffer(raw[y * (w * ch + 1) + 1:(y + 1) * (w * ch + 1)], np.uint8).astype(np.int32)
        cur = np.zeros_like(row)
        for i in range(len(row)):
            a = cur[i - ch] if i >= ch else 0
            b = prev[i]
            c = prev[i - ch] if i >= ch else 0
            if f == 0:
                p = 0
            elif f == 1:
                p = a
            elif f == 2:
                p = b
            elif f == 3:
                p = (a + b) // 2
            else:
                q = a + b - c
                pa, pb, pc = abs(q - a), abs(q - b), abs(q - c)
                p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
            cur[i] = (row[i] + p) & 0xFF
        out[y] = cur
        prev = cur
    if depth == 16:                                      # samples big-endian
        b = out.reshape(h, w * ch // 2, 2)
        return (b[..., 0] * 256 + b[..., 1]).reshape(h, w, ch // 2).astype(np.uint16), kinds
    return out.reshape(h, w, ch).astype(np.uint8), kinds


def png_text(data):
    """The iTXt chunks of a PNG: {keyword: text}, each read as the specification lays it out (keyword NUL, compression
    flag, method, language tag NUL, translated keyword NUL, UTF-8 text)."""
    pos, out = 8, {}
    while pos < len(data):
        n, = struct.unpack('>I', data[pos:pos + 4])
        kind, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        if kind == b'iTXt':
            key, rest = body.split(b'\0', 1)
            assert rest[0] == 0 and rest[1] == 0                 # uncompressed
            lang, rest = rest[2:].split(b'\0', 1)
            _, text = rest.split(b'\0', 1)
            out[key.decode('latin1')] = text.decode('utf8')
        pos += 12 + n
    return out


@pytest.mark.parametrize('filters', ['adaptive', 0, 1, 2, 3, 4])
def test_png_round_trip(filters):
    rng = np.random.default_rng(3)
    img = np.zeros((17, 23, 4), np.uint8)
    img[..., 0] = np.arange(23)[None, :] * 11          # gradients and noise, so every filter has work
    img[..., 1] = np.arange(17)[:, None] * 15
    img[..., 2] = rng.integers(0, 256, (17, 23))
    img[..., 3] = 255
    back, kinds = decode_png(png.encode(img, filters=filters))
    assert np.array_equal(back, img) and b'sRGB' in kinds


def test_javascript_png_decodes_the_same(tmp_path):
    rng = np.random.default_rng(4)
    img = rng.integers(0, 256, (9, 13, 4)).astype(np.uint8)
    src = tmp_path / 'img.bin'
    src.write_bytes(img.tobytes())
    dst = tmp_path / 'out.png'
    js = os.path.join(ROOT, 'shadytracer_amd', 'js', 'png.js')
    subprocess.run(['node', '-e', "const {encodePNG} = require(%s); const fs = require('fs');"
                    "fs.writeFileSync(%s, encodePNG(new Uint8Array(fs.readFileSync(%s)), 13, 9))" % (
                        json.dumps(js), json.dumps(str(dst)), json.dumps(str(src)))], check=True)
    back, _ = decode_png(dst.read_bytes())
    assert np.array_equal(back, img)
