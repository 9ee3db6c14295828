"""The light files (shadytracer_amd/js/export.js), CPU only: a seeded synthetic accumulation written as PFM, Radiance HDR
and OpenEXR by the JS writers, read back here by decoders written from the formats' published descriptions (not from
the writers), and the writers' known answers from the reference's own export test
(lib/pathtracer/test/export.test.mjs:229-261: rgbe(1, 0.5, 0.25) is 128,64,32,129; toHalf of 1, 65504 and 65520; the
ZIP pre-pass of a flat run)."""
import json
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
W, H = 37, 21


@pytest.fixture(scope='module')
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp('light')
    r = subprocess.run(['node', os.path.join(ROOT, 'tests', 'js', 'export_files.js'), str(d), str(W), str(H)],
                       check=True, stdout=subprocess.PIPE, text=True)
    acc = np.frombuffer((d / 'acc.bin').read_bytes(), np.float64).reshape(H, W, 4)
    # the rows a light file holds: top first (the accumulation is bottom first), 0 where no sample landed, float32
    want = np.where(acc[..., 3:] > 0, acc[..., :3], 0.0)[::-1].astype(np.float32)
    return d, json.loads(r.stdout), want


def test_known_answers(files):
    _, known, _ = files
    assert known['rgbe'][0] == [128, 64, 32, 129] and known['rgbe'][1] == [0, 0, 0, 0] and known['rgbe'][2] == [0, 0, 0, 0]
    assert known['rgbe'][3] == [255, 128, 0, 128]
    assert known['half'][:3] == [0x3c00, 0x7bff, 0x7c00]
    # against numpy's float16 (round to nearest even; subnormals; infinity; NaN's quiet bit)
    vals = [1, 65504, 65520, -2, 0.5, 6.103515625e-05, 5.960464477539063e-08, 2.98e-08, 1e-9, 1 + 2 ** -11, 1 + 3 * 2 ** -11, np.inf]
    with np.errstate(over='ignore'):
        assert known['half'][:len(vals)] == [int(np.array(v, np.float32).astype(np.float16).view(np.uint16)) for v in vals]
    assert known['half'][-1] == 0x7e00
    assert known['prepassFlat'][0] == 7 and set(known['prepassFlat'][1:]) == {128}


def test_pfm_holds_every_value_bottom_row_first(files):
    d, _, want = files
    b = (d / 'light.pfm').read_bytes()
    head = b.split(b'\n', 3)
    assert head[0] == b'PF' and head[1] == b'%d %d' % (W, H) and float(head[2]) == -1.0
    data = np.frombuffer(head[3], '<f4').reshape(H, W, 3)[::-1]
    assert np.array_equal(data, want)


def _rgbe_decode(p):
    m, e = p[..., :3].astype(np.float64), p[..., 3].astype(np.int64)
    return np.where(e[..., None] > 0, (m + 0.5) * np.ldexp(1.0, (e - 136))[..., None], 0.0)


def _read_hdr(b):
    lines, o = [], 0
    while True:
        n = b.index(b'\n', o)
        line = b[o:n].decode()
        o = n + 1
        if line == '':
            break
        lines.append(line)
    n = b.index(b'\n', o)
    res = b[o:n].decode().split()
    o = n + 1
    h, w = int(res[1]), int(res[3])
    assert res[0] == '-Y' and res[2] == '+X'
    px, rle = np.zeros((h, w, 4), np.uint8), 0
    for y in range(h):
        if 8 <= w <= 0x7fff:
            assert b[o] == 2 and b[o + 1] == 2 and (b[o + 2] << 8 | b[o + 3]) == w
            o += 4
            rle += 1
            for c in range(4):
                x = 0
                while x < w:
                    n = b[o]
                    o += 1
                    if n > 128:
                        px[y, x:x + n - 128, c] = b[o]
                        o += 1
                        x += n - 128
                    else:
                        px[y, x:x + n, c] = np.frombuffer(b[o:o + n], np.uint8)
                        o += n
                        x += n
                assert x == w
        else:
            px[y] = np.frombuffer(b[o:o + 4 * w], np.uint8).reshape(w, 4)
            o += 4 * w
    return lines, px, rle, o == len(b)


def test_hdr_in_radiance_units_within_the_rgbe_step(files):
    d, _, want = files
    lines, px, rle, end = _read_hdr((d / 'light.hdr').read_bytes())
    assert lines[0] == '#?RADIANCE' and 'FORMAT=32-bit_rle_rgbe' in lines and 'EXPOSURE=2.0' in lines
    assert any(l.startswith('PRIMARIES=0.640 0.330') for l in lines)
    assert rle == H and end
    v = np.maximum(want.astype(np.float64), 0) / 179          # RGBE holds no negative: clipped
    got = _rgbe_decode(px)
    step = v.max(-1, keepdims=True) / 128
    assert (np.abs(got - v) <= step + 1e-30).all()
    # a row narrower than 8 is written flat
    lines, px, rle, end = _read_hdr((d / 'narrow.hdr').read_bytes())
    assert rle == 0 and end and px.shape == (2, 5, 4)


def _read_exr(b):
    magic, version = struct.unpack_from('<ii', b, 0)
    assert magic == 20000630 and version == 2
    o, attrs = 8, {}
    while b[o] != 0:
        n = b.index(b'\0', o)
        name = b[o:n].decode()
        t = b.index(b'\0', n + 1)
        typ = b[n + 1:t].decode()
        size, = struct.unpack_from('<i', b, t + 1)
        attrs[name] = (typ, b[t + 5:t + 5 + size])
        o = t + 5 + size
    o += 1
    chl, chans = attrs['channels'][1], []
    p = 0
    while chl[p] != 0:
        n = chl.index(b'\0', p)
        ptype, = struct.unpack_from('<i', chl, n + 1)
        chans.append((chl[p:n].decode(), ptype))
        p = n + 17
    x0, y0, x1, y1 = struct.unpack('<4i', attrs['dataWindow'][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    comp = attrs['compression'][1][0]
    per = 16 if comp == 3 else 1
    nblocks = (h + per - 1) // per
    offsets = struct.unpack_from('<%dQ' % nblocks, b, o)
    bpv = 2 if chans[0][1] == 1 else 4
    img, zips = np.zeros((h, w, 3)), 0
    for off in offsets:
        y, size = struct.unpack_from('<ii', b, off)
        k = min(per, h - y)
        raw_len = k * w * 3 * bpv
        data = b[off + 8:off + 8 + size]
        if comp == 3 and size < raw_len:
            zips += 1
            t = bytearray(zlib.decompress(data))
            for i in range(1, len(t)):
                t[i] = (t[i - 1] + t[i] - 128) & 0xff
            half = (len(t) + 1) // 2
            raw = bytearray(len(t))
            raw[0::2], raw[1::2] = t[:half], t[half:]
            data = bytes(raw)
        assert len(data) == raw_len
        vals = np.frombuffer(data, '<f2' if bpv == 2 else '<f4').astype(np.float64).reshape(k, 3, w)
        for ci, (name, _) in enumerate(chans):
            img[y:y + k, :, 'RGB'.index(name)] = vals[:, ci, :]
    chroma = struct.unpack('<8f', attrs['chromaticities'][1])
    white, = struct.unpack('<f', attrs['whiteLuminance'][1])
    strings = {k: v[1].decode() for k, v in attrs.items() if v[0] == 'string'}
    return dict(img=img, chans=chans, comp=comp, zips=zips, chroma=chroma, white=white, strings=strings, order=attrs['lineOrder'][1][0])


def test_exr_float_zip_gives_every_value_back(files):
    d, _, want = files
    r = _read_exr((d / 'light.exr').read_bytes())
    assert [c for c, _ in r['chans']] == ['B', 'G', 'R'] and all(t == 2 for _, t in r['chans'])
    assert r['comp'] == 3 and r['zips'] >= 1 and r['order'] == 0
    assert np.array_equal(r['img'].astype(np.float32), want)            # the sun and the negatives included
    assert np.allclose(r['chroma'], [0.64, 0.33, 0.30, 0.60, 0.15, 0.06, 0.3127, 0.3290], atol=1e-7) and r['white'] == 1
    assert r['strings']['software'] == 'shadytracer_amd test'


def test_exr_half_none_within_half_rounding(files):
    d, _, want = files
    r = _read_exr((d / 'half.exr').read_bytes())
    assert all(t == 1 for _, t in r['chans']) and r['comp'] == 0
    exposed = (want.astype(np.float64) * 1e-3).astype(np.float32)
    with np.errstate(over='ignore'):
        assert np.array_equal(r['img'], exposed.astype(np.float16).astype(np.float64))   # past 65504: infinity
    assert np.isinf(r['img']).any() and abs(r['white'] - 1000) < 1e-3
