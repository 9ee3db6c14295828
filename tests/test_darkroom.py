"""The darkroom: its FP64 checker (oracle/develop.py develop_pixels and its stages) and its host mirror
(shadytracer_amd/js/darkroom.js: the white balance, the eyedropper, which stages a setting turns on, the curves'
table, the wheels) against what fixes them independently -- the reference's published hash values, Oklab's defining
properties, a white balance that must make its white (1, 1, 1), Fritsch-Carlson's monotonicity, the closed forms of
Levels and the CDL, the vignette's centre and corners -- and the C ABI's refusals of settings it cannot develop."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

import develop as D

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope='module')
def js():
    out = subprocess.run(['node', os.path.join(ROOT, 'tests', 'js', 'darkroom_values.js')], capture_output=True, text=True,
                         timeout=120, check=True)
    return json.loads(out.stdout)


# ---- the checker's stages

def test_pcg_hash_published_values():
    # the pcg_hash of Jarzynski and Olano (2020), as the reference's grade tests hold it: 0 -> 129708002, 1 -> 2831084092
    assert D.pcg_hash([0, 1]).tolist() == [129708002, 2831084092]
    assert D.unit24(0xFFFFFFFF) == 16777215 / 16777216 and D.unit24(255) == 0


def test_gamut_clip_keeps_luminance_and_zeroes_the_negative_channel():
    c = np.array([[1.0, -0.1, 0.3], [0.05, 0.4, -0.2], [0.2, 0.3, 0.4], [-1.0, 0.1, 0.1]])
    g = D.gamut_clip(c)
    assert np.allclose(g[:2] @ D.LUMA, c[:2] @ D.LUMA) and np.allclose(g[:2].min(axis=1), 0.0)
    # the hue kept as far as the gamut allows: the clipped colour lies on the line from the colour to its grey
    for k in range(2):
        Y = c[k] @ D.LUMA
        s = (g[k] - Y) / (c[k] - Y)
        assert np.allclose(s, s[0])
    assert np.array_equal(g[2], c[2]) and np.array_equal(g[3], [0, 0, 0])


def test_oklab_properties():
    for c in ([0.2, 0.5, 0.9], [3, 0.01, 0.4], [1.2, -0.1, 0.3], [1e-4, 2e-4, 5e-5]):
        assert np.allclose(D.from_oklab(D.oklab(c)), c, rtol=1e-9, atol=1e-12)
    assert np.allclose(D.oklab([1, 1, 1]), [1, 0, 0], atol=1e-6)
    # Ottosson's own value for linear sRGB red
    assert np.allclose(D.oklab([1, 0, 0]), [0.627955, 0.224863, 0.125846], atol=2e-6)
    # lightness scales as the cube root of the light
    assert D.oklab([8, 8, 8])[0] == pytest.approx(2 * D.oklab([1, 1, 1])[0])
    c = np.array([0.6, 0.2, 0.1])
    L, a, b = D.oklab(c)
    assert np.allclose(D.lab_adjust(c, 0, 0, 0), c)
    s = D.oklab(D.lab_adjust(c, 0.5, 0, 0))
    assert s[0] == pytest.approx(L) and np.hypot(s[1], s[2]) == pytest.approx(1.5 * np.hypot(a, b))
    assert np.arctan2(s[2], s[1]) == pytest.approx(np.arctan2(b, a))
    h = D.oklab(D.lab_adjust(c, 0, 0, np.pi / 2))
    assert h[0] == pytest.approx(L) and np.arctan2(h[2], h[1]) - np.arctan2(b, a) == pytest.approx(np.pi / 2)

    def gain(x, k):
        return np.hypot(*D.oklab(D.lab_adjust(np.array(x), 0, k, 0))[1:]) / np.hypot(*D.oklab(x)[1:])
    dull, vivid = [0.5, 0.48, 0.46], [0.9, 0.1, 0.05]
    assert gain(dull, 0.5) > 1.4 and gain(vivid, 0.5) == pytest.approx(1.0)
    assert gain([v * 1000 for v in dull], 0.5) == pytest.approx(gain(dull, 0.5))


def test_equaliser_bands_interpolated_and_held():
    b = [1, 2, 3, 4, 5, 6, 7, 8]
    assert D.eq_gain(b, -7) == 1 and D.eq_gain(b, -3) == 5 and D.eq_gain(b, 0) == 8
    assert D.eq_gain(b, -2.25) == pytest.approx(5.75)
    assert D.eq_gain(b, -20) == 1 and D.eq_gain(b, 3) == 8


def test_curve7_at_its_defaults_is_hables_curve():
    x = np.array([0, 0.01, 0.18, 1, 4, 5.6])
    assert np.allclose(D.tone7(x, 5.6, 0.20, 0.15), D.tone_hable(x), rtol=1e-12, atol=1e-15)
    assert D.tone7(3.0, 3.0, 0.3, 0.4) == pytest.approx(1.0)


def test_finish_closed_forms():
    x = np.array([0, 0.001, 0.18, 0.5, 1])
    assert np.allclose(D.encode_out(x, 0), D.srgb(x)) and np.allclose(D.encode_out(x, 1), x ** (1 / 2.2))
    assert np.allclose(D.encode_out(x, 2), x ** (1 / 2.4)) and np.array_equal(D.encode_out(x, 3), x)
    assert D.encode_out(2.0, 1) == 1 and D.encode_out(-1.0, 0) == 0
    assert np.allclose(D.levels(np.array([0, 0.3, 1]), 0, 1, 1, 0, 1), [0, 0.3, 1])
    assert D.levels(0.1, 0.1, 0.9, 1, 0, 1) == 0 and D.levels(0.95, 0.1, 0.9, 1, 0, 1) == 1
    assert D.levels(0.25, 0, 1, 0.5, 0, 1) == pytest.approx(0.5)                  # gamma 2 (1/gamma 0.5) lifts the middle
    assert D.levels(0.5, 0, 1, 1, 0.1, 0.9) == pytest.approx(0.5) and D.levels(0, 0, 1, 1, 0.1, 0.9) == pytest.approx(0.1)
    assert D.cdl(0.5, 2, 0.1, 0.5) == pytest.approx(np.sqrt(1.1)) and D.cdl(0.1, 1, -0.5, 1) == 0
    lut = np.concatenate([np.arange(1024) / 1023.0] * 4)
    xs = np.linspace(0, 1, 37)
    assert np.allclose(D.lut_at(lut, 2, xs), xs)


def test_vignette_centre_and_corners():
    assert D.vignette_stops(0, 0, 100, 50, 0, 0.5, 0) == 0
    assert D.vignette_stops(49.5, 24.5, 100, 50, -1, 0.3, 0) == 0
    assert D.vignette_stops(-0.5, -0.5, 100, 50, -1, 0.3, 0) == pytest.approx(-1)
    assert D.vignette_stops(60, 25, 100, 50, -1, 0.5, 0) == 0
    assert D.vignette_stops(0, 24.5, 100, 50, -1, 0.3, 1) < D.vignette_stops(49.5, 0, 100, 50, -1, 0.3, 1)


def test_grain_and_dither_statistics():
    y, x = np.mgrid[0:64, 0:64].astype(float)
    g = D.grain_noise(x, y, 1.5)
    assert g.min() >= -1 and g.max() <= 1 and abs(g.mean()) < 0.1
    # value noise: on a cell's corner pixel centre it is that cell's own value; smooth in between
    assert np.all(np.abs(np.diff(D.grain_noise(np.arange(200.0), np.zeros(200), 40.0))) < 0.1)
    y, x = np.mgrid[0:100, 0:100].astype(float)
    d = D.dither_at(x, y)
    assert d.min() >= -1 / 255 and d.max() <= 1 / 255 and abs(d.mean()) < 0.05 / 255
    # the sum of two uniforms: triangular, variance 2 / 12 of a step squared
    assert (d * 255).var() == pytest.approx(1 / 6, rel=0.05)


def test_turbo_runs_from_dark_blue_to_dark_red():
    # Turbo's published ends: a dark blue-violet at 0 (its table's first entry 0.190, 0.072, 0.232), green-yellow in the
    # middle, dark red at 1 (0.480, 0.016, 0.011); the fit is within 0.05 of the ends' brightness order
    blue, mid, red = D.turbo([0.1, 0.5, 0.9])
    assert blue[2] > blue[1] > blue[0] and red[0] > red[1] > red[2] and mid[1] > max(mid[0], mid[2])
    assert np.argmax(D.turbo([0.0])[0]) != 1 and np.argmax(D.turbo([1.0])[0]) == 0
    t = np.linspace(0, 1, 256)
    assert np.all((D.turbo(t) >= -0.03) & (D.turbo(t) <= 1.035))     # the fit overshoots a little; the view clamps


def test_identity_grade_is_the_plain_develop():
    rng = np.random.default_rng(3)
    acc = np.concatenate([rng.lognormal(0, 2, (24, 32, 3)), np.ones((24, 32, 1))], axis=-1)
    for tone in range(7):
        rgb, v, ev = D.develop(acc, tone=tone)
        want = D.srgb(D.TONES[tone](D.gamut_clip(acc[..., :3] / (D.EV_CLIP * 2.0 ** ev)))) * 255
        assert np.allclose(v, want[::-1])
    # the histogram counts the developed values
    h = D.histogram(acc, tone=1)
    assert h.sum(axis=1).tolist() == [24 * 32] * 4
    assert np.array_equal(h[0], np.bincount(D.develop(acc, tone=1)[0][..., 0].ravel(), minlength=256))


def test_false_colour_reads_the_number():
    acc = np.zeros((40, 10, 4))
    acc[..., :3] = 100.0 / np.pi                        # E / pi on a white: 100 lux
    acc[..., 3] = 1
    g = dict(measureMode=1, measureMin=1.0, measureMax=10000.0)
    _, v, _ = D.develop(acc, tone=1, ev=0.0, grade=g)
    assert np.allclose(v[0, 0] / 255, D.turbo(0.5))     # log2 100 is half way from log2 1 to log2 10^4 (top row)
    assert np.allclose(v[-1, 3] / 255, np.clip(D.turbo(3.5 / 10), 0, 1))   # the bottom row is the legend strip
    _, v2, _ = D.develop(acc, tone=1, ev=0.0, grade=dict(g, measureMode=2))
    assert np.allclose(v2[0, 0] / 255, D.turbo(np.log2(100 / np.pi) / np.log2(1e4)))
    _, v3, _ = D.develop(acc, tone=1, ev=0.0, grade=dict(g, measureBands=4))
    assert np.allclose(v3[0, 0] / 255, D.turbo(0.625))


# ---- the host mirror

def test_white_balance_defaults_are_rec709s_white(js):
    assert js['wbDefault']['temp'] == 6506 and js['wbDefault']['duv'] == pytest.approx(0.0033, abs=1e-4)
    assert js['d65'] == pytest.approx([0.312716, 0.329001], abs=1e-6)
    # Krystek's locus at 2856 K against CIE illuminant A (0.44757, 0.40745), within its stated accuracy
    assert js['illuminantA'] == pytest.approx([0.44757, 0.40745], abs=5e-4)


def test_white_balance_makes_its_white_white(js):
    for k, out in js['white'].items():
        assert out == pytest.approx([1, 1, 1], abs=1e-9), k
    x, y = js['moved']
    assert x < 0.40 and y < 0.45                         # a 3000 K balance moves other colours bluer
    assert js['gained'][0] == pytest.approx(2 * js['plain'][0]) and js['gained'][2] == pytest.approx(0.5 * js['plain'][2])


def test_eyedropper(js):
    e = js['eyedropper']
    assert e['temp'] == 3000 and abs(e['tint'] + js['wbDefault']['duv'] * 1000) <= 0.05
    w = js['balancedWall']
    assert max(w) / min(w) < 1.003
    assert all(js['refused'])


def test_which_stages_a_setting_turns_on(js):
    b = js['bits']
    ANY, WB, EQ, LAB, LV, CU, CDL, VIG, GR, DI, LUMA = 1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024
    assert b['defaults'] == 0 and b['empty'] == 0
    assert b['wbTemp'] == b['wbTint'] == b['gainB'] == ANY | WB
    assert b['eq3'] == ANY | EQ and b['labVib'] == ANY | LAB and b['curve7'] == ANY
    assert b['otherCurves'] == [0] * 7
    assert b['levels'] == ANY | LV and b['curves'] == ANY | CU and b['curvesLuma'] == ANY | CU | LUMA and b['lumaAlone'] == 0
    assert b['cdl'] == ANY | CDL and b['vignette'] == ANY | VIG and b['grain'] == ANY | GR and b['dither'] == ANY | DI
    assert b['transfer'] == ANY
    u = js['uniforms']
    assert u['wb'] == pytest.approx(u['wantWb']) and u['eq'][0] == 1 and u['eq'][7] == -1
    assert u['lab'][2] == pytest.approx(np.pi / 2) and u['tone7'][0] == pytest.approx(5.6) and u['levels'][2] == 0.5


def test_curves_are_monotone_cubics_through_their_points(js):
    c = js['curve']
    assert c['at'] == pytest.approx([p[1] for p in c['points']])
    xs, ys = np.array(c['xs']), np.array(c['ys'])
    # monotone data stays monotone, and between two points the curve stays within them (Fritsch-Carlson)
    assert np.all(np.diff(ys) >= -1e-12)
    pts = np.array(c['points'])
    for (x0, y0), (x1, y1) in zip(pts[:-1], pts[1:]):
        seg = ys[(xs >= x0) & (xs <= x1)]
        assert seg.min() >= min(y0, y1) - 1e-12 and seg.max() <= max(y0, y1) + 1e-12
    flat = ys[(xs >= 0.6) & (xs <= 0.8)]
    assert np.all(flat == pytest.approx(0.62))
    assert np.all(ys[xs <= 0.0] == 0.1)
    # a straight curve is i / 1023 exactly (in float32); a bent master leaves the colour curves straight
    straight = np.float32(np.arange(1024) / 1023.0)
    assert np.array_equal(np.float32(js['straightLut']), np.tile(straight, 4))
    bent = np.float32(js['bentLut'])
    assert bent[512] > straight[512] and np.array_equal(bent[1024:], np.tile(straight, 3))
    assert js['normalized'] == [[0, 0.2], [0.5, 1], [1, 1]]


def test_wheels(js):
    c = js['cdl']
    assert c['rest']['offset'] == [0, 0, 0] and c['rest']['power'] == [1, 1, 1] and c['rest']['slope'] == [1, 1, 1]
    assert sum(c['tint']) == pytest.approx(0, abs=1e-12)
    assert c['red'][0] > 0 and c['red'][1] < 0 and c['red'][2] < 0
    assert c['lift']['offset'] == pytest.approx([0.05] * 3) and c['lift']['slope'] == pytest.approx([2] * 3)
    assert js['step'] == [2, 6, 1]


# ---- the C ABI's refusals (no GPU call: each is refused before anything reaches the device)

def test_develop_refuses_what_it_cannot_develop():
    from shadytracer_amd import pt
    lib = pt.load_library()
    d = pt.PtDevelop()
    hist = (ctypes.c_uint32 * 1024)()
    # no context
    assert lib.pt_develop(None, ctypes.byref(d), ctypes.sizeof(d), None, 0, None) != 0
    assert lib.pt_develop_histogram(None, ctypes.byref(d), ctypes.sizeof(d), 1, hist, 1024, None) != 0
    assert ctypes.sizeof(pt.PtDevelop) == lib.pt_struct_size(15) == 64 * 4 + 4 * 1024 * 4


def test_views_of_the_checker():
    rng = np.random.default_rng(5)
    acc = np.concatenate([rng.lognormal(0, 1.5, (20, 32, 3)), np.ones((20, 32, 1))], axis=-1)
    rgb, v, ev = D.develop(acc, tone=1)
    same, sv = D.develop_view(acc, (16.0, 10.0, 1.0, 1.0), 32, 20, tone=1, ev=ev)
    assert np.array_equal(same, rgb) and np.allclose(sv, v)
    # a half-size thumbnail is the mean of each 2 x 2 block of the picture's values
    _, tv = D.develop_view(acc, (16.0, 10.0, 2.0, 2.0), 16, 10, tone=1, ev=ev)
    assert np.allclose(tv, v.reshape(10, 2, 16, 2, 3).mean(axis=(1, 3)))
    # zoomed in 4 x: each picture pixel a 4 x 4 square; beyond the picture the backdrop
    _, zv = D.develop_view(acc, (30.0, 18.0, 0.25, 0.25), 24, 24, tone=1, ev=ev)
    assert np.allclose(zv[:, -4:], 0.09 * 255) and np.allclose(zv[:4, :], 0.09 * 255)     # x >= 32 or y >= 20 is outside
    assert np.allclose(zv[-1, 0], v[20 - 1 - 15, 27])     # the bottom-left output pixel looks at picture pixel (27, 15)
    assert np.allclose(zv[-4:, :4], v[20 - 1 - 15, 27])   # ... and so do the other 15 of its 4 x 4 square


def test_recipes_travel_and_are_read_with_care(js):
    r = js['recipes']
    assert r['recipe']['kind'] == 'darkroom-recipe' and r['recipe']['settings']['curves']['m'] == [[0, 0], [0.5, 0.6], [1, 1]]
    assert r['back']['settings'] == r['recipe']['settings'] and r['back']['ignored'] == []
    # a number as text is read as the number; a boolean where a number belongs, a number where a boolean belongs and an
    # unknown key are left out; a curve that is not one is straight
    assert r['odd']['settings']['wbTemp'] == 3200 and 'labVib' not in r['odd']['settings'] and 'dither' not in r['odd']['settings']
    assert sorted(r['odd']['ignored']) == ['dither', 'labVib', 'mystery']
    assert r['odd']['settings']['curves']['m'] == [[0, 0], [1, 1]]
    assert r['fromPicture']['settings']['wbTemp'] == 3400 and r['fromPicture']['settings']['ev'] == 15
    assert 'error' in r['notRecipe'] and 'error' in r['notJson'] and 'error' in r['bare']
