"""The developed picture's files from JavaScript (js/jpeg.js, js/png.js) without a GPU: the JPEG against an
independent decoder (PIL's libjpeg) and libjpeg's own quantisation tables at the same quality, and the 8- and 16-bit
PNGs with their iTXt text against the reader written from the specification in test_develop."""
import io
import json
import os
import subprocess

import numpy as np
import pytest

import test_develop as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
Image = pytest.importorskip('PIL.Image')


@pytest.fixture(scope='module')
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp('pictures')
    raw = d / 'src.raw'
    r = subprocess.run(['node', os.path.join(ROOT, 'tests', 'js', 'picture_files.js'), str(raw), str(d)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    meta = json.loads(r.stdout)
    src = np.fromfile(str(raw), np.uint8).reshape(meta['h'], meta['w'], 4)
    return d, src


def psnr(a, b):
    return 10 * np.log10(255.0 ** 2 / np.mean((a.astype(float) - b.astype(float)) ** 2))


@pytest.mark.parametrize('q, floor', [(92, 38.0), (50, 30.0), (100, 45.0)])
def test_jpeg_decodes_to_the_picture(files, q, floor):
    d, src = files
    im = Image.open(d / ('q%d.jpg' % q))
    assert im.format == 'JPEG' and im.mode == 'RGB' and im.size == (src.shape[1], src.shape[0])
    got = np.asarray(im)
    assert psnr(got, src[..., :3]) > floor
    # libjpeg at the same quality: the same tables (IJG's scaling of Annex K), and pictures alike
    buf = io.BytesIO()
    Image.fromarray(src[..., :3]).save(buf, 'JPEG', quality=q, subsampling=2)
    ref = Image.open(io.BytesIO(buf.getvalue()))
    assert {k: list(v) for k, v in im.quantization.items()} == {k: list(v) for k, v in ref.quantization.items()}
    assert np.abs(np.asarray(ref).astype(float) - got).mean() < 1.5
    assert im.layer == ref.layer                        # 4:2:0: luma sampled 2 x 2, chroma 1 x 1


def test_png_text_and_16_bits(files):
    d, src = files
    for name in ('p8.png', 'p8late.png'):
        data = (d / name).read_bytes()
        img, kinds = T.decode_png(data)
        assert np.array_equal(img, src) and kinds[0] == b'IHDR' and kinds[1] == b'iTXt'
        text = T.png_text(data)
        assert text['Software'] == 'shadytracer_amd' and json.loads(text['Comment']) == {'unit': 'cd/m²', 'ev': 9.5}
    data = (d / 'p16.png').read_bytes()
    img, kinds = T.decode_png(data)
    want = np.fromfile(str(d / 'p16.raw'), np.uint16).reshape(src.shape[0], src.shape[1], 3)
    assert img.dtype == np.uint16 and np.array_equal(img, want)
    assert json.loads(T.png_text(data)['Comment'])['unit'] == 'cd/m²'
    # and PIL reads it as a 16-bit picture too
    im = Image.open(io.BytesIO(data))
    im.load()
    assert im.size == (src.shape[1], src.shape[0]) and im.info.get('Software') == 'shadytracer_amd'


@pytest.mark.parametrize('channels', [3, 4])
def test_python_png_16_bits_with_text(channels):
    from shadytracer_amd import png
    rng = np.random.default_rng(11)
    img = rng.integers(0, 65536, (13, 19, channels)).astype(np.uint16)
    img[..., 0] = np.arange(19)[None, :] * 3000                  # a gradient, so the filters have work
    data = png.encode(img, text=[('Software', 'shadytracer_amd'), ('Comment', 'ev 9.5, cd/m²')])
    got, kinds = T.decode_png(data)
    assert got.dtype == np.uint16 and np.array_equal(got, img) and kinds[1] == b'iTXt'
    assert T.png_text(data) == {'Software': 'shadytracer_amd', 'Comment': 'ev 9.5, cd/m²'}
    for k in range(5):                                           # every filter on 6- and 8-byte pixels
        assert np.array_equal(T.decode_png(png.encode(img, filters=k))[0], img)
