"""The optics trace over the scene on the GPU (pt_optics_run through the C ABI): the reference's own optics rows
(lib/pathtracer/tests.js optics-balance :1060-1100, optics-farfield :1101-1137, optics-refraction :1138-1200; expected
values and the reference's measured ones from its fixture, tests/golden/reference_rows.json) and the closed forms they
stand for. Each ray is traced through the engine's own march, surface kernel and materials."""
import json
import math
import os
import sys

import numpy as np
import pytest

from shadytracer_amd import farfield, pt, scenes

pytestmark = pytest.mark.gpu

ROWS = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'reference_rows.json')))['tests']


def row(test, case):
    for r in ROWS[test]['rows']:
        if r['case'] == case:
            return r
    raise KeyError(case)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def load(engine, name, **opts):
    made = scenes.emit(name, **opts)
    engine.set_scene(made['scene'])
    engine.set_uniforms(pt.uniforms_for(made))
    return made


def mat(made, ident):
    """The engine's material id of the document's material `ident` (id 0 is the emitter's default material)."""
    return [m['id'] for m in made['scene']['materials']].index(ident)


def fresnel(cos_i, n1, n2):
    """Unpolarised Fresnel reflectance (the mean of s and p)."""
    s2 = (n1 / n2) ** 2 * (1 - cos_i ** 2)
    if s2 >= 1:
        return 1.0
    ct = math.sqrt(1 - s2)
    rs = ((n1 * cos_i - n2 * ct) / (n1 * cos_i + n2 * ct)) ** 2
    rp = ((n1 * ct - n2 * cos_i) / (n1 * ct + n2 * cos_i)) ** 2
    return 0.5 * (rs + rp)


DOWN = {'type': 'rectangle', 'pos': (0, 0, 1), 'axis': (0, 0, -1), 'u': (1, 0, 0), 'size': (0.5, 0.5),
        'distribution': 'cosine', 'halfAngle': 90, 'n': 1, 'flux': 1000}


@pytest.mark.parametrize('case,opts,source,key,value', [
    ('floor albedo 0.8', {'albedo': 0.8}, {'distribution': 'collimated'}, 'absorbed', 0.2),
    ('floor albedo 0.5', {'albedo': 0.5}, {'distribution': 'collimated'}, 'absorbed', 0.5),
    ('black floor', {'albedo': 0.0}, {'distribution': 'collimated'}, 'absorbed', 1.0),
    ('mirror floor', {'albedo': 1.0, 'metallic': 1, 'roughness': 0, 'specular': 1}, {'distribution': 'collimated'}, 'escaped', 1.0),
    ('albedo 0.5, Lambertian source', {'albedo': 0.5}, {}, 'absorbed', 0.5),
])
def test_balance_on_a_floor(engine, case, opts, source, key, value):
    # a source 1 m over a floor shining down: the floor takes 1 - albedo, the rest leaves the scene; the tallies
    # close exactly (emitted = escaped + lost + volume + sum of incident - leaving)
    say('case: optics balance, %s' % case)
    r = row('optics-balance', case)
    assert r['expected'] == pytest.approx(value)
    load(engine, 'optics-floor', **opts)
    res = engine.optics_run([dict(DOWN, **source)], rays=1 << 20, bounces=64, seed=7)
    share = res[key] / res['emitted']
    say('  %s share %.6f (expected %g, the reference measured %g)' % (key, share, r['expected'], r['measured']))
    assert share == pytest.approx(r['expected'], abs=4e-3)
    taken = sum(m['incident'] - m['leaving'] for m in res['materials'])
    assert res['escaped'] + res['lost'] + res['volume'] + taken == pytest.approx(res['emitted'], rel=1e-6)
    assert res['rays'] == 1 << 20


def test_max_distance_and_a_detector_floor(engine):
    say('case: optics balance, max distance 0.5 m and the floor a detector')
    made = load(engine, 'optics-floor', albedo=0.5)
    floor = mat(made, 'floor')
    # straight down: every ray would meet the floor 1 m away (a Lambertian source also sends a few grazing rays out of
    # the scene before they reach it)
    down = dict(DOWN, distribution='collimated')
    res = engine.optics_run([down], rays=1 << 18, seed=3, max_dist=0.5)
    assert res['lost'] / res['emitted'] == pytest.approx(row('optics-balance', 'MAX DISTANCE 0.5 m, floor 1 m away: none arrive')['expected'])
    assert res['materials'][floor]['hits'] == 0
    res = engine.optics_run([DOWN], rays=1 << 18, seed=3, max_dist=0.5)
    assert res['lost'] / res['emitted'] > 0.997 and res['lost'] + res['escaped'] == pytest.approx(res['emitted'])
    res = engine.optics_run([down], rays=1 << 18, seed=3, detectors=[floor], det_cap=4096)
    assert res['detected'] / res['emitted'] == pytest.approx(row('optics-balance', 'the floor a DETECTOR: every ray caught')['expected'])
    assert res['detectorHits'] == 1 << 18
    h = res['hits']
    # the hits: on the floor, under the source's square, arriving downwards, each with the flux one ray carries
    assert h.shape == (4096, 8)
    assert np.all(np.abs(h[:, 2]) < 1e-3) and np.all(np.abs(h[:, :2]) <= 0.5 + 1e-3)
    assert np.all(h[:, 6] < 0)
    assert np.allclose(h[:, 3], 1000.0 / (1 << 18), rtol=1e-5)
    assert np.all(h[:, 7].view(np.uint32) == floor)


@pytest.mark.parametrize('albedo', [0.8, 0.5])
def test_integrating_sphere(engine, albedo):
    # a point source inside a closed Lambertian shell: each ray meets the wall 1 / (1 - albedo) times on average
    case = 'integrating sphere, albedo %g' % albedo
    say('case: optics balance, %s' % case)
    made = load(engine, 'optics-shell', albedo=albedo)
    src = {'type': 'point', 'pos': (0, 0, 0), 'axis': (0, 0, 1), 'distribution': 'uniform', 'halfAngle': 90, 'flux': 1000}
    # a path of many bounces takes many steps near the walls: the budget is raised past what 64 bounces need
    res = engine.optics_run([src], rays=1 << 18, bounces=64, seed=11, march_steps=30000)
    arrived = res['materials'][mat(made, 'wall')]['incident'] / res['emitted']
    expect = row('optics-balance', case)['expected']
    cap = (1 - albedo ** 65) / (1 - albedo)      # 64 bounces: what the series reaches
    say('  arrived %.5f (expected %g; %.5f within 64 bounces)' % (arrived, expect, cap))
    assert arrived == pytest.approx(cap, rel=5e-3)
    assert abs(arrived - expect) / expect < 1e-2
    assert res['escaped'] == 0.0


def far_trace(engine, source, rays=1 << 21, far=None, scene_opts=None, detect_floor=False):
    made = load(engine, 'optics-floor', **(scene_opts or {'albedo': 0.5}))
    floor = mat(made, 'floor')
    detectors = [floor] if detect_floor else []
    if far and far.get('mode') == 'detector':
        far = dict(far, material=floor)
    src = {'type': 'disc', 'pos': (0, 0, 1), 'axis': (0, 0, 1), 'u': (1, 0, 0), 'size': (0.1, 0.1),
           'distribution': 'cosine', 'halfAngle': 90, 'n': 1, 'flux': 1000}
    src.update(source)
    f = {'mode': 'escaped', 'nC': 24, 'nG': 37, 'axis': (0, 0, 1), 'c0': (1, 0, 0)}
    f.update(far or {})
    return engine.optics_run([src], rays=rays, bounces=8, seed=5, far=f, detectors=detectors)


def test_far_field_of_a_lambertian_disc(engine):
    say('case: optics far field, Lambertian disc shining up')
    res = far_trace(engine, {}, rays=1 << 22)
    F = res['far']
    assert F['flux'].sum() / res['emitted'] == pytest.approx(row('optics-farfield', 'Lambertian: everything reaches the far field')['expected'], abs=1e-6)
    I = farfield.intensity(F, 'rotational')
    i0 = farfield.i0_fit(I) / (1000 / math.pi)
    beam = farfield.width(I, 0.0, 0.5)
    r1, r2 = row('optics-farfield', 'Lambertian: I(0) = flux / pi'), row('optics-farfield', 'Lambertian: beam angle 120 deg')
    say('  I0 ratio %.5f (reference measured %.5f), beam %.3f (reference measured %.3f)' % (i0, r1['measured'], beam, r2['measured']))
    assert i0 == pytest.approx(r1['expected'], abs=5e-3)
    assert beam == pytest.approx(r2['expected'], rel=0.02)
    # the far field binned by the grid's definition: every row's intensity on cos gamma (the Lambertian law) within
    # its Poisson noise, up to 85 degrees
    g = np.radians(np.arange(37) * 5.0)
    Ir = I.mean(axis=1)
    for j in range(1, 18):
        assert Ir[j] == pytest.approx(1000 / math.pi * math.cos(g[j]), rel=0.03), j


def grid_far(share, nC=24, nG=37):
    """The far field a source sends on the grid exactly: each gamma row's share of the flux, from the cumulative
    share(gamma) of the flux within gamma of the axis, spread evenly over the row's C bins."""
    dG = 180.0 / (nG - 1)
    F = np.zeros((nG, nC))
    for j in range(nG):
        lo, hi = math.radians(max(0.0, (j - 0.5) * dG)), math.radians(min(180.0, (j + 0.5) * dG))
        F[j, :] = 1000.0 * (share(hi) - share(lo)) / nC
    return {'flux': F, 'nC': nC, 'nG': nG}


@pytest.mark.parametrize('case,source,closed,share', [
    # cos^n in solid angle: the flux within gamma is 1 - cos^(n+1) gamma
    ('cos^5: beam angle', {'n': 5}, 2 * math.degrees(math.acos(0.5 ** (1 / 5))),
     lambda g: 1.0 - max(math.cos(min(g, math.pi / 2)), 0.0) ** 6),
    # uniform within 30 degrees: (1 - cos gamma) / (1 - cos 30), capped at 1
    ('uniform 30-deg cone: beam angle 60 deg', {'distribution': 'uniform', 'halfAngle': 30}, 60.0,
     lambda g: min((1.0 - math.cos(g)) / (1.0 - math.cos(math.radians(30))), 1.0)),
])
def test_beam_angles(engine, case, source, closed, share):
    say('case: optics far field, %s' % case)
    r = row('optics-farfield', case)
    assert r['expected'] == pytest.approx(closed, abs=1e-4)
    res = far_trace(engine, source, rays=1 << 23)
    I = farfield.intensity(res['far'], 'rotational')
    beam = farfield.width(I, 0.0, 0.5)
    # the same estimator on the exact far field of this source, binned on the same 5-degree grid: what the trace must
    # read but for its noise (reading the half level between bin centres moves it off the closed form)
    on_grid = farfield.width(farfield.intensity(grid_far(share), 'rotational'), 0.0, 0.5)
    say('  beam %.4f (exact on this grid %.4f, closed form %.4f; the reference measured %.4f on the same grid)' % (
        beam, on_grid, closed, r['measured']))
    # (8 M rays: the cap bin that sets the half level holds ~60 K of them, 0.4 % noise, ~0.1 degree of width)
    assert beam == pytest.approx(on_grid, abs=0.4)
    # the reference's row passes within 2 % of the closed form, and its own run read 58.25 and 59.62
    assert beam == pytest.approx(closed, rel=0.02)
    assert abs(beam - r['measured']) < 1.0


def test_far_field_fed_by_a_detector(engine):
    say('case: optics far field fed by a detector floor lit from above')
    res = far_trace(engine, {'axis': (0, 0, -1)}, far={'mode': 'detector', 'axis': (0, 0, -1)}, detect_floor=True)
    I = farfield.intensity(res['far'], 'rotational')
    i0 = farfield.i0_fit(I) / (1000 / math.pi)
    r = row('optics-farfield', 'fed by a DETECTOR (the floor, lit from above): I(0) = flux / pi')
    say('  I0 ratio %.5f (reference measured %.5f)' % (i0, r['measured']))
    assert i0 == pytest.approx(r['expected'], abs=5e-3)
    # all but the grazing rays that leave the scene before reaching the floor (0.25 % of a Lambertian source 20 m out)
    assert 0.997 < res['detected'] / res['emitted'] <= 1.0


@pytest.mark.parametrize('deg', [0, 45, 70])
def test_slab_transmission(engine, deg):
    # a collimated beam through a smooth glass slab (n 1.5) onto a black detector floor: T = (1 - R) / (1 + R) with
    # every internal reflection summed, R the unpolarised Fresnel at the incidence angle
    case = 'slab, %d deg' % deg
    say('case: optics refraction, %s' % case)
    made = load(engine, 'optics-slab')
    th = math.radians(deg)
    src = {'type': 'disc', 'pos': (0, 0, 1), 'axis': (math.sin(th), 0, -math.cos(th)), 'u': (0, 1, 0), 'size': (0.3, 0.3),
           'distribution': 'collimated', 'flux': 1000}
    res = engine.optics_run([src], rays=1 << 20, bounces=64, seed=13, detectors=[mat(made, 'floor')])
    T = res['detected'] / res['emitted']
    R = fresnel(math.cos(th), 1.0, 1.5)
    expect = row('optics-refraction', case)['expected']
    say('  T %.5f (expected %.5f, closed form %.5f)' % (T, expect, (1 - R) / (1 + R)))
    assert expect == pytest.approx((1 - R) / (1 + R), rel=1e-6)
    assert T == pytest.approx(expect, abs=3e-3)


def test_absorbing_slab_takes_its_share(engine):
    # the same slab head-on, its glass absorbing to 0.5 at 50 mm (thickness 100 mm: tau = 0.25 inside): the detector
    # takes (1 - R)^2 tau / (1 - R^2 tau^2), and the glass what its inside took, the balance whole
    say('case: optics absorbing slab, 0 deg')
    made = load(engine, 'optics-slab', attenuation=0.05)
    src = {'type': 'disc', 'pos': (0, 0, 1), 'axis': (0, 0, -1), 'u': (0, 1, 0), 'size': (0.3, 0.3),
           'distribution': 'collimated', 'flux': 1000}
    res = engine.optics_run([src], rays=1 << 20, bounces=64, seed=23, detectors=[mat(made, 'floor')])
    R, tau = 0.04, 0.25
    T = res['detected'] / res['emitted']
    want = (1 - R) ** 2 * tau / (1 - R * R * tau * tau)
    back = R + (1 - R) ** 2 * R * tau * tau / (1 - R * R * tau * tau)
    say('  T %.5f (closed form %.5f), escaped %.5f (closed form %.5f)' % (T, want, res['escaped'] / res['emitted'], back))
    assert T == pytest.approx(want, abs=2e-3)
    assert res['escaped'] / res['emitted'] == pytest.approx(back, abs=2e-3)
    took = res['materials'][mat(made, 'glass')]['absorbed']
    assert took / res['emitted'] == pytest.approx(1 - want - back, abs=3e-3)
    assert res['emitted'] == pytest.approx(res['escaped'] + res['detected'] + res['absorbed'] + res['lost'], rel=1e-6)


def test_glass_ball_interactions(engine):
    # a point source at the centre of a glass ball (emitting into n 1.5): every ray meets the surface head-on, 4 % comes
    # back through the centre to try again -- 1 / (1 - R) interactions a ray, and every ray escapes
    say('case: optics refraction, glass ball with the source at its centre')
    load(engine, 'optics-ball')
    src = {'type': 'point', 'pos': (0, 0, 0), 'axis': (0, 0, 1), 'distribution': 'uniform', 'halfAngle': 90, 'flux': 1000,
           'emitN': 1.5, 'twoSided': True}
    res = engine.optics_run([src], rays=1 << 20, bounces=64, seed=17)
    per = res['interactions'] / res['rays']
    expect = row('optics-refraction', 'glass ball, source at its centre')['expected']
    say('  interactions a ray %.5f (expected %.5f)' % (per, expect))
    assert per == pytest.approx(expect, abs=2e-3)
    assert res['escaped'] / res['emitted'] == pytest.approx(1.0, abs=1e-4)


def test_same_seed_same_rays(engine):
    # the trace is deterministic: the tallies are fixed point and the draws keyed by ray id and seed
    say('case: optics trace repeatability')
    load(engine, 'optics-floor', albedo=0.5)
    a = engine.optics_run([DOWN], rays=100000, seed=21)
    b = engine.optics_run([DOWN], rays=100000, seed=21)
    c = engine.optics_run([DOWN], rays=100000, seed=22)
    assert a['escaped'] == b['escaped'] and a['materials'] == b['materials']
    assert a['escaped'] != c['escaped']


def test_bad_arguments_are_refused(engine):
    say('case: optics trace refusals')
    load(engine, 'optics-floor', albedo=0.5)
    with pytest.raises(pt.PtError, match='no sources'):
        engine.optics_run([dict(DOWN, flux=0)], rays=1000)
    with pytest.raises(pt.PtError, match='far field'):
        engine.optics_run([DOWN], rays=1000, far={'mode': 'escaped', 'nG': 1})
    with pytest.raises(pt.PtError, match='bounces'):
        engine.optics_run([DOWN], rays=1000, bounces=300)


def test_node_create_optics_runs_the_trace():
    # the JavaScript surface (createOptics(RT): setSource, setFar, run, readHits) over the addon's worker thread
    say('case: node tests/js/optics_trace.js, createOptics over the optics floor')
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run(['node', os.path.join(root, 'tests', 'js', 'optics_trace.js')], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout)
    say('  %s' % {k: js[k] for k in ('escaped', 'beam', 'field', 'ticks', 'detected', 'hitCount')})
    assert js['batches'] == 4 and js['rays'] == 4 << 20
    assert js['escaped'] == pytest.approx(1000.0) and js['farSum'] == pytest.approx(1000.0, rel=1e-9)
    assert abs(js['balance']) < 1e-6 * js['emitted']
    assert js['beam'] == pytest.approx(120.0, rel=0.02)
    # a Lambertian source: I falls to a tenth at acos(0.1), 84.26 degrees
    assert js['field'] == pytest.approx(2 * math.degrees(math.acos(0.1)), rel=0.02)
    assert js['ticks'] > 0
    assert js['detected'] == pytest.approx(500.0) and js['hitCount'] == 1 << 16 and js['hitsKept'] == 1000
    h = js['firstHit']
    assert abs(h['pos'][2]) < 1e-3 and h['dir'][2] < -0.999 and h['flux'] == pytest.approx(500.0 / (1 << 16), rel=1e-5)
    # the run gives its seed back: the same run again traces the same rays
    assert js['seed'] == 1



def test_detector_map_of_a_lambertian_source(engine):
    # a small Lambertian disc 1 m over a detector floor, facing it: E(r) = flux h^2 / (pi (h^2 + r^2)^2), 318.3 lx under
    # it; the hits mapped (shadytracer_amd/detmap.py) in 5 cm cells over 2 x 2 m, ring by ring against the closed form
    from shadytracer_amd import detmap
    say('case: optics detector map under a Lambertian disc')
    made = load(engine, 'optics-floor', albedo=0.5)
    floor = mat(made, 'floor')
    src = {'type': 'disc', 'pos': (0, 0, 1), 'axis': (0, 0, -1), 'size': (0.01, 0.01), 'distribution': 'cosine', 'flux': 1000}
    n = 1 << 22
    res = engine.optics_run([src], rays=n, seed=23, detectors=[floor], det_cap=n)
    assert res['detectorHits'] > 0.99 * n and len(res['hits']) == res['detectorHits']
    m = detmap.detector_map(res['hits'], cells=40, rng={'a': (-1.0, 1.0), 'b': (-1.0, 1.0)})
    assert (m['nx'], m['ny']) == (40, 40) and abs(m['normal'][2]) == 1.0
    ax = m['range']['a'][0] + (np.arange(40) + 0.5) * m['cell'] + m['centre'][0]
    ay = m['range']['b'][0] + (np.arange(40) + 0.5) * m['cell'] + m['centre'][1]
    X, Y = np.meshgrid(ax, ay)
    r2 = X * X + Y * Y
    closed = 1000.0 / (math.pi * (1.0 + r2) ** 2)
    for lo, hi in ((0.0, 0.15), (0.15, 0.3), (0.3, 0.5), (0.5, 0.7), (0.7, 0.9)):
        ring = (r2 >= lo * lo) & (r2 < hi * hi)
        ratio = m['E'][ring].sum() / closed[ring].sum()
        say('  r %.2f-%.2f m: E / closed form %.4f over %d cells' % (lo, hi, ratio, ring.sum()))
        assert ratio == pytest.approx(1.0, abs=0.02)
