"""The developed picture on the GPU (pt_develop through the C ABI) against the FP64 checker (oracle/develop.py) on the
same accumulation: the meter's exposure, every tone curve, the sRGB encoding and the 8-bit rounding; and the picture
saved as PNG from JavaScript."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import develop as D
from shadytracer_amd import png, pt, scenes

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def render(engine, name, samples, **opts):
    made = scenes.emit(name, **opts)
    engine.set_scene(made['scene'])
    engine.set_uniforms(pt.uniforms_for(made))
    engine.clear_accums()
    engine.trace(samples)
    engine.sync()
    return engine.read_accum()[0]


def test_develop_matches_the_oracle(engine):
    say('case: develop room-c2 96x54 (16 samples), every tone curve, metered and set exposure')
    acc = render(engine, 'room-c2', 16, width=96, height=54)
    for tone in range(7):
        rgba, info = engine.develop(tone=tone, auto=True, speed=1.0, headroom=2.0)
        want, v, ev = D.develop(acc, tone=tone, headroom=2.0)
        assert info['ev'] == pytest.approx(ev, abs=2e-3), (tone, info['ev'], ev)
        assert info['metered'] == (acc[..., 3] > 0).sum()
        diff = np.abs(rgba[..., :3].astype(int) - want.astype(int))
        say('  tone %d: EV %.4f (oracle %.4f), %d of %d channels off by 1, none by more' % (
            tone, info['ev'], ev, int((diff == 1).sum()), diff.size))
        # FP32 on the GPU against FP64: a value within rounding of a half step may round either way
        assert diff.max() <= 1 and (diff > 0).mean() < 0.01, tone
        assert np.all(rgba[..., 3] == 255)
    rgba, info = engine.develop(tone='aces', auto=False, ev=9.0)
    want, _, _ = D.develop(acc, tone=4, ev=9.0)
    assert info['ev'] == 9.0 and np.abs(rgba[..., :3].astype(int) - want.astype(int)).max() <= 1
    # rows: top first with flip, as accumulated without
    flat, _ = engine.develop(tone='aces', auto=False, ev=9.0, flip=False)
    assert np.array_equal(flat[::-1], rgba)


def test_auto_exposure_eases_and_starts_afresh(engine):
    say('case: develop plane-c1 64x64, the meter easing and its reset')
    acc = render(engine, 'plane-c1', 4, width=64, height=64)
    _, a = engine.develop(auto=True, speed=1.0)
    assert a['ev'] == pytest.approx(D.develop(acc)[2], abs=2e-3)
    # held: speed 0 keeps the exposure whatever the picture says
    _, b = engine.develop(auto=True, speed=0.0, headroom=-5.0)
    assert b['ev'] == a['ev']
    # a brighter sky, a new picture: the first development after clearAccums jumps to the new reading
    u = pt.uniforms_for(scenes.emit('plane-c1', width=64, height=64))
    u.env.L *= 16.0
    engine.set_uniforms(u)
    engine.clear_accums()
    engine.trace(4)
    engine.sync()
    _, c = engine.develop(auto=True, speed=0.1)
    assert c['ev'] == pytest.approx(a['ev'] + 4.0, abs=0.35)


def test_node_saves_a_png(tmp_path):
    say('case: node tests/js/develop_png.js, plane-c1 through createPathTracer, savePicture')
    out = tmp_path / 'c1.png'
    accf = tmp_path / 'acc.f64'
    out16, jpg = tmp_path / 'c1-16.png', tmp_path / 'c1.jpg'
    r = subprocess.run(['node', os.path.join(ROOT, 'tests', 'js', 'develop_png.js'), str(out), str(accf), str(out16), str(jpg)],
                       stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout)
    import test_develop as T
    img, kinds = T.decode_png(out.read_bytes())
    assert img.shape == (js['height'], js['width'], 4) and b'sRGB' in kinds
    acc = np.fromfile(str(accf), np.float64).reshape(js['height'], js['width'], 4)
    want, v, ev = D.develop(acc, tone=1)
    assert js['ev'] == pytest.approx(ev, abs=2e-3)
    assert np.abs(img[..., :3].astype(int) - want.astype(int)).max() <= 1
    # the 16-bit file: the same picture to 1/65535 (FP32 stages against FP64), saying how it was made
    data16 = out16.read_bytes()
    img16, _ = T.decode_png(data16)
    assert img16.dtype == np.uint16 and img16.shape == (js['height'], js['width'], 3)
    assert np.abs(img16.astype(float) - v * 257.0).max() < 0.002 * 65535
    assert json.loads(T.png_text(data16)['Comment'])['format'] == 'png16'
    # the JPEG: a JFIF file that decodes close to the picture
    assert jpg.read_bytes()[:4] == b'\xff\xd8\xff\xe0'
    try:
        from PIL import Image
    except ImportError:
        return
    got = np.asarray(Image.open(str(jpg)).convert('RGB')).astype(float)
    assert np.abs(got - want).mean() < 3.0


# ---- the darkroom

def darkroom_words(settings):
    r = subprocess.run(['node', os.path.join(ROOT, 'tests', 'js', 'darkroom_uniforms.js'), json.dumps(settings)], capture_output=True,
                       text=True, timeout=60, check=True)
    return json.loads(r.stdout)


BEND = {'m': [[0, 0], [0.3, 0.36], [1, 1]], 'r': [[0, 0.02], [0.5, 0.47], [1, 1]]}
GRADES = {
    'light': dict(wbTemp=3400, wbTint=3, gainB=1.05, eq1=0.5, eq5=-0.4, labSat=0.3, labVib=0.4, labHue=12),
    'finish': dict(transfer=1, lvInBlack=0.02, lvInWhite=0.95, lvGamma=1.2, lvOutBlack=0.01, lvOutWhite=0.98, curves=BEND,
                   liftX=0.1, liftM=0.05, gammaY=-0.2, gainM=0.1),
    'luma-curve': dict(curves=BEND, curveLuma=True, vigAmount=-0.8, vigMid=0.2, vigRound=0.5),
    'curve7-graded': dict(toneMap=7, tcWhite=2.0, tcToe=0.35, tcShoulder=0.25, grainAmount=0.5, grainSize=2.0, dither=True),
}


@pytest.mark.parametrize('name', list(GRADES))
def test_darkroom_matches_the_checker(engine, name):
    settings = GRADES[name]
    say('case: develop room-c2 96x54 (16 samples), darkroom', name)
    acc = render(engine, 'room-c2', 16, width=96, height=54)
    grade = darkroom_words(settings)
    if name == 'finish':
        grade.update(contrast=0.15, saturation=-0.1, brightness=0.02)
    tone = settings.get('toneMap', 1)
    rgba, info = engine.develop(tone=tone, auto=False, ev=9.0, grade=grade)
    want, v, _ = D.develop(acc, tone=tone, ev=9.0, grade=grade)
    diff = np.abs(rgba[..., :3].astype(int) - want.astype(int))
    # FP32 against FP64: a channel lands one 8-bit step off only where the exact value sits at a rounding edge
    say('  max step', diff.max(), 'channels off', int((diff > 0).sum()), 'of', diff.size)
    assert diff.max() <= 1 and (diff > 0).mean() < 0.01
    # the stages did something: the graded picture is not the plain one
    plain, _, _ = D.develop(acc, tone=tone if tone != 7 else 6, ev=9.0)
    assert np.abs(want.astype(int) - plain.astype(int)).mean() > 1.0


def test_develop16_matches_the_checker(engine):
    say('case: develop16 room-c2 96x54 (16 samples), finish grade, dither asked and left off')
    acc = render(engine, 'room-c2', 16, width=96, height=54)
    grade = darkroom_words(dict(GRADES['finish'], dither=True))
    rgba16, _ = engine.develop16(tone=1, auto=False, ev=9.0, grade=grade)
    rgba8, _ = engine.develop(tone=1, auto=False, ev=9.0, grade=dict(grade, lightOn=grade['lightOn'] & ~512))
    _, v, _ = D.develop(acc, tone=1, ev=9.0, grade=dict(grade, lightOn=grade['lightOn'] & ~512))
    assert rgba16.dtype == np.uint16 and np.all(rgba16[..., 3] == 65535)
    err = np.abs(rgba16[..., :3].astype(float) - v * 257.0)
    say('  16-bit error max', err.max(), 'mean', err.mean())
    assert err.max() < 0.002 * 65535 and err.mean() < 8
    # the 8-bit picture is the 16-bit one rounded, but at rounding edges
    assert np.abs(np.round(rgba16[..., :3] / 257.0) - rgba8[..., :3]).max() <= 1


def ambiguous_taps(view, w, h):
    """[h, w] (top row first): output pixels with a tap within 1e-3 of a picture pixel's edge, where an FP32 tap and
    the checker's FP64 one may read neighbouring pixels."""
    cx0, cy0, sx, sy = view
    box = int(min(8, max(1, np.ceil(max(sx, sy) - 1e-9))))
    oy, ox = np.mgrid[0:h, 0:w].astype(float)
    bad = np.zeros((h, w), bool)
    for b in range(box * box):
        qx = cx0 + (ox + 0.5 - 0.5 * w) * sx - 0.5 * sx + ((b % box) + 0.5) / box * sx
        qy = cy0 + (oy + 0.5 - 0.5 * h) * sy - 0.5 * sy + ((b // box) + 0.5) / box * sy
        bad |= (np.abs(qx - np.round(qx)) < 1e-3) | (np.abs(qy - np.round(qy)) < 1e-3)
    return bad[::-1]


def test_views_of_the_developed_picture(engine):
    say('case: develop_view room-c2 96x54 (16 samples): full size, thumbnail, zoomed past the edge, clip stripes')
    acc = render(engine, 'room-c2', 16, width=96, height=54)
    grade = darkroom_words(GRADES['light'])
    full, _ = engine.develop(tone=1, auto=False, ev=9.0, grade=grade)
    same, _ = engine.develop_view((48.0, 27.0, 1.0, 1.0), 96, 54, tone=1, auto=False, ev=9.0, grade=grade)
    assert np.array_equal(same, full)                  # the whole picture at its own size is the development itself
    for view, w, h, clip, ev in (((48.0, 27.0, 2.0, 2.0), 48, 27, False, 9.0),          # a thumbnail: 2 x 2 taps
                                 ((48.3, 27.1, 3.1337, 3.1337), 30, 17, False, 9.0),    # an odd one: 4 x 4 taps
                                 ((94.0, 2.0, 0.25, 0.25), 64, 48, False, 9.0),         # zoomed in across the corner
                                 ((48.0, 27.0, 1.0, 1.0), 96, 54, True, 6.0)):          # over-exposed, stripes
        got, _ = engine.develop_view(view, w, h, clip_warn=clip, tone=1, auto=False, ev=ev, grade=grade)
        want, _ = D.develop_view(acc, view, w, h, tone=1, ev=ev, grade=grade, clip_warn=clip)
        sure = ~ambiguous_taps(view, w, h)
        assert sure.mean() > 0.8, view
        diff = np.abs(got[..., :3].astype(int) - want.astype(int))[sure]
        say('  view', view, w, h, 'max step', diff.max(), 'off', int((diff > 0).sum()), 'compared', int(sure.sum()))
        assert diff.max() <= 1 and (diff > 0).mean() < 0.01, view
    # the backdrop beyond the picture's corner, and stripes where an over-exposed picture clips
    corner, _ = engine.develop_view((94.0, 2.0, 0.25, 0.25), 64, 48, tone=1, auto=False, ev=9.0, grade=grade)
    assert np.all(corner[:, -1, :3] == 23) and np.all(corner[-1, :, :3] == 23)
    striped, _ = engine.develop_view((48.0, 27.0, 1.0, 1.0), 96, 54, clip_warn=True, tone=1, auto=False, ev=6.0, grade=grade)
    assert (np.all(striped[..., :3] == [255, 0, 0], axis=-1)).sum() > 0


def test_false_colour_views_match_the_checker(engine):
    say('case: develop room-c2 96x54 (16 samples), false colour of luminance, continuous and in 6 bands')
    acc = render(engine, 'room-c2', 16, width=96, height=54)
    for g in (dict(measureMode=2, measureMin=0.5, measureMax=5000.0), dict(measureMode=2, measureMin=0.5, measureMax=5000.0, measureBands=6)):
        rgba, _ = engine.develop(tone=1, auto=False, ev=9.0, grade=g)
        want, _, _ = D.develop(acc, tone=1, ev=9.0, grade=g)
        diff = np.abs(rgba[..., :3].astype(int) - want.astype(int))
        assert diff.max() <= 1 and (diff > 0).mean() < 0.01
        # the legend strip along the bottom: the scale itself, left to right
        assert np.array_equal(rgba[-1, :, :3], want[-1])


def test_histogram_counts_the_developed_picture(engine):
    say('case: develop_histogram room-c2 96x54 (16 samples), step 1 and 3, graded')
    acc = render(engine, 'room-c2', 16, width=96, height=54)
    grade = darkroom_words(GRADES['light'])
    rgba, _ = engine.develop(tone=1, auto=False, ev=9.0, grade=grade, flip=False)
    for step in (1, 3):
        hist, _ = engine.develop_histogram(tone=1, auto=False, ev=9.0, grade=grade, step=step)
        sub = rgba[::step, ::step, :3]
        for ch in range(3):
            assert np.array_equal(hist[ch], np.bincount(sub[..., ch].ravel(), minlength=256)), (step, ch)
        assert int(hist[3].sum()) == sub.shape[0] * sub.shape[1]
        want = D.histogram(acc, tone=1, ev=9.0, grade=grade, step=step)
        assert np.abs(hist[3].astype(int) - want[3]).sum() <= 0.02 * sub.shape[0] * sub.shape[1]
    with pytest.raises(pt.PtError, match='false-colour'):
        engine.develop_histogram(tone=1, grade=dict(measureMode=1))
