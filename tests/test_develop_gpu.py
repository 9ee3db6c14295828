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
    r = subprocess.run(['node', os.path.join(ROOT, 'tests', 'js', 'develop_png.js'), str(out), str(accf)], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout)
    import test_develop as T
    img, kinds = T.decode_png(out.read_bytes())
    assert img.shape == (js['height'], js['width'], 4) and b'sRGB' in kinds
    acc = np.fromfile(str(accf), np.float64).reshape(js['height'], js['width'], 4)
    want, _, ev = D.develop(acc, tone=1)
    assert js['ev'] == pytest.approx(ev, abs=2e-3)
    assert np.abs(img[..., :3].astype(int) - want.astype(int)).max() <= 1
