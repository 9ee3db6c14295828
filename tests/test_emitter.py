"""The scene emitter's rules (shadytracer_amd/js/hipgen.js), CPU only: a light's colour against the reference's own
colorimetry expectations (lib/pathtracer/core/test/colorimetry.test.mjs:29-66: the tabulated blackbody locus, D65,
a royal blue outside Rec.709), the document's orbit view as the engine's camera, the studio environment, and what the
emitter refuses rather than draws wrongly."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SCENES = '/root/reference/scenes'


@pytest.fixture(scope='module')
def rules():
    r = subprocess.run(['node', os.path.join(ROOT, 'tests', 'js', 'emitter_rules.js')], check=True, stdout=subprocess.PIPE, text=True)
    return json.loads(r.stdout)


M = np.array([[0.4124, 0.3576, 0.1805], [0.2126, 0.7152, 0.0722], [0.0193, 0.1192, 0.9505]])   # IEC 61966-2-1


def test_planckian_locus_against_the_tabulated_blackbody(rules):
    # colorimetry.test.mjs:31-35: the CIE 2-degree blackbody xy, within 5e-4
    table = {2000: (0.5267, 0.4133), 2856: (0.4476, 0.4074), 3000: (0.4369, 0.4041), 4000: (0.3805, 0.3768),
             5000: (0.3451, 0.3516), 6500: (0.3135, 0.3236), 10000: (0.2807, 0.2884)}
    worst = max(math.hypot(x - table[T][0], y - table[T][1]) for T, x, y in rules['locus'])
    assert worst < 5e-4, worst


def test_duv_is_a_perpendicular_step_towards_green(rules):
    step, above = rules['duvStep']
    assert abs(step - 0.01) < 1e-9 and above


def test_light_colours_have_luminance_one_and_their_chromaticity(rules):
    assert np.allclose(rules['d65'], 1, atol=2e-3)                 # D65 is Rec.709 white
    blue = np.array(rules['blue'])
    assert blue.min() < 0                                           # outside the gamut: kept, not clipped
    XYZ = M @ blue
    assert abs(XYZ[1] - 1) < 1e-12
    assert np.allclose(XYZ[:2] / XYZ.sum(), [0.15, 0.025], atol=1e-12)
    # illuminant A's chromaticity (0.44757, 0.40745) from 2856 K, to the locus fit's 5e-4
    XYZ = M @ np.array(rules['cct'])
    assert math.hypot(*(XYZ[:2] / XYZ.sum() - [0.44757, 0.40745])) < 5e-4
    # an RGB colour is only normalised to luminance 1
    assert np.allclose(np.array(rules['rgb']) / rules['rgb'][0], [1, 0.5, 0]) and abs(M[1] @ rules['rgb'] - 1) < 1e-12


def test_a_lights_flux_is_pi_L_A(rules):
    # 100 lm from a 1 m^2 one-sided Lambertian face: L = 100 / pi (lights.js fluxFactor: k = pi)
    assert abs(M[1] @ rules['lightLe'] - 100 / math.pi) < 1e-9


def test_studio_puts_its_illuminance_on_a_level_surface(rules):
    # environment.js studioLight: L(d) = L0 (1 + top d.z); the upper hemisphere's cosine integral of it is the set
    # illuminance (here by quadrature in cos theta), the tint normalised to luminance 1
    e = rules['studioEnv']
    assert e['type'] == 3 and abs(e['top'] - 0.6) < 1e-12
    mu = (np.arange(200000) + 0.5) / 200000
    E = 2 * math.pi * np.mean(e['L'] * (1 + e['top'] * mu) * mu)
    assert abs(E / 300 - 1) < 1e-9
    assert abs(M[1] @ e['col'] - 1) < 1e-12 and np.allclose(np.array(e['col']) / e['col'][0], [1, 0.5, 0.25])


def test_orbit_view_becomes_the_pinhole(rules):
    # lib/camera.js syncOrbitPose: target + distance (cos el cos az, cos el sin az, sin el), el = 90 - elevation
    c = rules['camera']
    el, az = math.radians(15), math.radians(60)
    want = np.array([0, 0, 0.12]) + 0.45 * np.array([math.cos(el) * math.cos(az), math.cos(el) * math.sin(az), math.sin(el)])
    assert np.allclose(c['pos'], want, atol=1e-12)
    back = np.array(c['fwd'])
    assert np.allclose(back, (want - [0, 0, 0.12]) / 0.45, atol=1e-12)       # from the target to the eye
    r, u = np.array(c['right']), np.array(c['up'])
    assert abs(r[2]) < 1e-12 and abs(r @ u) < 1e-12 and np.allclose(np.cross(r, u), back)
    assert abs(c['fy'] - math.tan(math.radians(20))) < 1e-12 and abs(c['fx'] - 2 * c['fy']) < 1e-12


def test_what_the_engine_does_not_draw_is_refused_by_name(rules):
    R = rules['refusals']
    assert R['studio'] is None and R['comments'] is None and R['lambertian'] is None and R['unusedcoating'] is None
    assert R['sunsky'] is None and R['nosetting'] is None and "scene's own code" in R['envcode']
    assert "material 'm': its code: unknown function 'myTexture' (line 2" in R['code'] and R['translated'] is None
    assert 'coating' in R['coating'] and "its relief's code: unknown function 'heightAt'" in R['relief']
    assert 'collimated' in R['collimated'] and 'cone' in R['cone']
    assert rules['approximations'] == ["material 'm': anisotropy"]


from sky_quadrature import level_sky_lux as _level_sky_lux, uniforms_with as _uniforms_with   # noqa: E402


def _cct(rgb):
    # McCamy's cubic in the chromaticity of linear Rec.709 rgb (a few kelvin off the Planckian CCT here)
    X, Y, Z = M @ np.asarray(rgb, float)
    x, y = X / (X + Y + Z), Y / (X + Y + Z)
    n = (x - 0.3320) / (0.1858 - y)
    return 449 * n ** 3 + 3525 * n ** 2 + 6823.3 * n + 5520.33


def test_sun_and_sky_meet_the_reference_figures(rules):
    # environment.js:103-106, printed by builder/test/sky.test.mjs: 128 klx above the air, turbidity 0.7. The sun 45
    # degrees up gives 91 klx head-on and the sky 11.5 klx on a level surface, the sun about 4980 K and the zenith
    # about 9800 K; 15 degrees up 51 klx and 4080 K; 5 degrees up 11 klx and 2810 K. Set more than 3 degrees: no sky.
    import oracle
    e45, e15, e5, eset = rules['sunsky']
    head_on = lambda e: float(M[1] @ e['sunL']) * math.pi * e['sunSin2']      # noqa: E731
    assert e45['type'] == 4 and abs(head_on(e45) - 91e3) < 500 and abs(_level_sky_lux(e45) - 11.5e3) < 100
    assert abs(head_on(e15) - 51e3) < 500 and abs(head_on(e5) - 11e3) < 500
    # the colour temperatures: the bands' colours come from a fit to the CIE functions, within 1 % of the table's
    for e, want in ((e45, 4980), (e15, 4080), (e5, 2810)):
        assert abs(_cct(e['sunL']) - want) < 60, (_cct(e['sunL']), want)
    zen = oracle.env(_uniforms_with(e45), [[0, 0, 1]])[0]
    assert abs(_cct(zen) - 9800) < 300, _cct(zen)
    assert eset['sunE'] == 0 and oracle.env(_uniforms_with(eset), [[0, 0, 1]]).max() == 0
    # the sun's place: azimuth from +y towards +x (sdf.js sunDirection:72)
    el, az = math.radians(45), math.radians(30)
    assert np.allclose(e45['sunDir'], [math.cos(el) * math.sin(az), math.cos(el) * math.cos(az), math.sin(el)], atol=1e-12)


def test_sky_tables_agree_across_the_emitter_the_device_and_the_oracle(rules):
    # the device's literals are the emitter's table; the oracle's own table (in C, double) gives the emitter's sky
    import oracle
    with open(os.path.join(ROOT, 'shadytracer_amd', 'csrc', 'device', 'pt_sky.h')) as f:
        assert rules['skyHeader'] in f.read()
    u = _uniforms_with(dict(rules['sunsky'][0], sunE=0.0))
    dirs = np.array([[0, 0, 1], [0.6, 0, 0.8], [0, -0.96, 0.28], [0.3, 0.3, -0.905]])
    got = oracle.env(u, dirs)
    np.testing.assert_allclose(got, np.array(rules['skyJs']), rtol=1e-6)      # the oracle reads the sun's place as f32
    # luminance 1 over the sixteen bands, and each band's depths from its middle
    tot = np.sum([b['rgb'] for b in rules['skyBands']], 0)
    assert abs(M[1] @ tot - 1) < 1e-12
    for b in rules['skyBands']:
        um = b['nm'] / 1000
        assert abs(b['tauR'] - 0.008569 * um ** -4 * (1 + 0.0113 * um ** -2 + 0.00013 * um ** -4)) < 1e-15


CIE_TABLE = '/root/reference/lib/pathtracer/core/cie1931.js'


@pytest.mark.skipif(not os.path.isfile(CIE_TABLE), reason="the reference's CIE table is not here")
def test_sky_band_colours_against_the_cie_table(rules):
    # the bands' colours from the CIE 1931 table itself (the reference's data file, read as numbers), against the fit's:
    # within 0.01 band by band (the largest shares are 0.2-0.5), and the sunlight's colour within 1 %
    import re
    src = open(CIE_TABLE).read()
    cmf = [np.array([float(x) for x in re.search(r'CIE_%s = Float64Array\.of\(([^)]*)\)' % c, src).group(1).split(',')]) for c in 'XYZ']
    planck = lambda nm: 1 / ((nm * 1e-9) ** 5 * (math.exp(1.438776877e-2 / (nm * 1e-9 * 5778)) - 1))   # noqa: E731
    Mi = np.linalg.inv(M)
    step, bands = 350 / 16, []
    for k in range(16):
        xyz, l = np.zeros(3), 380 + k * step
        while l < 380 + (k + 1) * step:
            i = int(round(l)) - 360
            xyz += planck(l) * np.array([cmf[0][i], cmf[1][i], cmf[2][i]])
            l += 1
        bands.append(Mi @ xyz)
    bands = np.array(bands) / (M[1] @ np.sum(bands, 0))
    fit = np.array([b['rgb'] for b in rules['skyBands']])
    assert np.abs(fit - bands).max() < 0.01, np.abs(fit - bands).max()
    assert np.allclose(fit.sum(0), bands.sum(0), rtol=0.01)


@pytest.mark.skipif(not os.path.isdir(REF_SCENES), reason='the reference scenes are not here')
def test_reference_scenes_emit_or_name_what_stops_them():
    # every scene document the reference ships either emits, or is refused with the node, material, light or
    # environment named -- never a crash or a silent substitute
    from shadytracer_amd import scenes
    ok = []
    for f in sorted(os.listdir(REF_SCENES)):
        if not f.endswith('.json'):
            continue
        try:
            made = scenes.emit_doc(os.path.join(REF_SCENES, f), width=64, height=36)
            ok.append(f)
            assert made['camera'] and made['scene']['approximations'] == []
        except RuntimeError as e:
            msg = str(e)
            assert 'the engine does not draw' in msg, (f, msg[-300:])
    # those with builtin shapes (lens, prism, reflector, aspheric lens, outlines and the solids among them), the studio
    # or the sun & sky, and Lambertian lights; eight more through their materials' own code, reliefs and absorption
    assert {'led-lens-screen.json', 'parametric-part.json', 'rgbw-mixing.json', 'dusty-beam.json', 'parabolic-reflector.json',
            'aspheric-condenser.json', 'fresnel-plano-convex.json', 'combining-shapes.json', 'shape-gallery.json',
            'depth-of-field.json', 'first-steps.json', 'lighting-room.json', 'office-lighting.json', 'relief-texture.json',
            'glass-liquid.json', 'product-shot.json', 'infinite-detail.json'} <= set(ok), ok
