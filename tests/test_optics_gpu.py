"""The optics bench on the GPU through the C ABI: the same photons as the FP64 oracle (PCG4D keyed alike), the numbers
recomputed from the GPU's own tallies by the oracle, closed forms, and the 64-scene persistent launch."""
import json
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from shadytracer_amd import optics, pt

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_numbers_from_the_gpus_own_tallies():
    # the FP64 numbers kernel against the oracle's, on the same tallies: they must agree to rounding
    say('case: optics numbers, 6 bench scenes at 2^16 photons, GPU tallies through both')
    sc = optics.bench_scenes(6, log2_min=16, log2_max=16)
    res, F, L, ms = pt.optics_trace(sc, far=True, lux=True)
    for s, r, f, l in zip(sc, res, F, L):
        o = oracle.optics_numbers(s, f, l, [r['X'], r['Y'], r['Z'], r['absorbed'], r['photons']])
        for k in ('Imax', 'gammaMax', 'x', 'y', 'duv', 'luxMax', 'luxCentre', 'efficiency'):
            assert o[k] == pytest.approx(r[k], rel=1e-9, abs=1e-12), (k, o[k], r[k])
        # CCT is where a squared distance is least: flat there to rounding, so two FP64 evaluations (the GPU's fused
        # multiply-adds, the CPU's separate ones) agree on it to about sqrt(eps) T, 1e-8 relative
        assert o['cct'] == pytest.approx(r['cct'], rel=1e-7), (o['cct'], r['cct'])
        for k in ('beam', 'field'):
            assert np.allclose(o[k], r[k], rtol=1e-9, atol=1e-9), (k, o[k], r[k])


def test_the_same_photons_as_the_oracle():
    # FP32 on the GPU, FP64 on the CPU, the same PCG4D draws: the far fields agree but for the rare photon whose bin or
    # Fresnel choice the rounding flips
    say('case: optics trace, 8 bench scenes at 2^17 photons, GPU vs FP64 oracle')
    sc = optics.bench_scenes(8, log2_min=17, log2_max=17, seed=99)
    g, GF, GL, _ = pt.optics_trace(sc, far=True, lux=True)
    o, OF, OL = oracle.optics_trace(sc, threads=0, tallies=True)
    for k, (a, b) in enumerate(zip(g, o)):
        assert a['photons'] == b['photons'] == sc[k].photons
        assert abs(a['escaped'] + a['absorbed'] - a['emitted']) < 1e-5 * a['emitted']
        assert abs(a['efficiency'] - b['efficiency']) < 2e-3, (k, a['efficiency'], b['efficiency'])
        assert abs(a['Imax'] / b['Imax'] - 1) < 0.02, (k, a['Imax'], b['Imax'])
        for q in ('beam', 'field'):
            assert abs(a[q][0] - b[q][0]) < max(0.1, 0.02 * b[q][0]), (k, q, a[q], b[q])
        assert abs(a['cct'] - b['cct']) < 1.0 and abs(a['duv'] - b['duv']) < 1e-4, (k, a['cct'], b['cct'])
        assert abs(a['luxCentre'] / b['luxCentre'] - 1) < 0.03, (k, a['luxCentre'], b['luxCentre'])
        # the walks' surface tests: the same photons take the same steps but where FP32 rounding flips a choice
        assert abs(a['events'] / b['events'] - 1) < 1e-3, (k, a['events'], b['events'])
        assert 40 < a['flops'] / a['events'] < 400, (k, a['flops'] / a['events'])
        # the far fields bin by bin: the flux that moved bins is a small share
        moved = np.abs(GF[k] - OF[k]).sum() / (2 * OF[k].sum())
        assert moved < 0.01, (k, moved)


def test_bare_led_closed_forms():
    say('case: optics, a bare Lambertian LED at 2^24 photons')
    res, F, L, _ = pt.optics_trace([optics.bare_led(1 << 24)], far=True, lux=True)
    r = res[0]
    dG = math.radians(1.0)
    I = np.array([F[0, j].sum() / (2 * math.pi * (math.cos(max(0, (j - 0.5) * dG)) - math.cos(min(math.pi, (j + 0.5) * dG))))
                  for j in range(181)])
    want = 100 / math.pi * np.maximum(np.cos(np.radians(np.arange(181))), 0)
    assert np.max(np.abs(I[3:85] / want[3:85] - 1)) < 0.02
    assert abs(r['efficiency'] - 1) < 1e-5
    assert abs(r['field'][0] - 2 * math.degrees(math.acos(0.1))) < 1.0 and 115 < r['beam'][0] < 121
    assert abs(L[0, 120:136, 120:136].mean() / (100 / math.pi) - 1) < 0.01


def test_sixty_four_scenes_in_one_launch():
    say('case: optics bench, 64 seeded luminaires, 2^18..2^24 photons each, one persistent launch')
    sc = optics.bench_scenes(64)
    res, _, _, (ms, ms_numbers) = pt.optics_trace(sc)
    total = sum(s.photons for s in sc)
    say('%d photons in %.1f ms (numbers %.3f ms): %.0f Mphotons/s' % (total, ms, ms_numbers, total / ms / 1e3))
    for s, r in zip(sc, res):
        assert r['photons'] == s.photons
        assert abs(r['escaped'] + r['absorbed'] - r['emitted']) < 1e-5 * r['emitted']
        assert 0.5 < r['efficiency'] <= 1.0 and r['Imax'] > 0 and 0 < r['beam'][0] < 40 and 2500 < r['cct'] < 7000


def test_node_addon_runs_the_optics_bench(tmp_path):
    # the JavaScript path (shadytracer_amd/js/optics.js -> addon -> pt_optics_trace) gives the ctypes path's numbers
    say('case: node tests/js/optics_run.js, 4 bench scenes at 2^18 photons, against the ctypes binding')
    sc = optics.bench_scenes(4, log2_min=18, log2_max=18, seed=7)
    doc = [{k: (list(getattr(s, k)) if k in ('xyz', 'rho') else getattr(s, k)) for k, _ in pt.PtOpticsScene._fields_} for s in sc]
    path = tmp_path / 'scenes.json'
    path.write_text(json.dumps(doc))
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run(['node', os.path.join(root, 'tests', 'js', 'optics_run.js'), str(path)], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout)
    py, F, _, _ = pt.optics_trace(sc, far=True)
    assert js['farSum'] == pytest.approx(F.sum(), rel=1e-12)
    for a, b in zip(js['results'], py):
        # per-chunk sums are added in a fixed order: the totals repeat exactly; the far-field numbers to rounding
        for k in ('photons', 'events', 'X', 'Y', 'Z', 'efficiency', 'cct', 'duv'):
            assert a[k] == b[k], (k, a[k], b[k])
        for k in ('Imax', 'luxMax', 'luxCentre'):
            assert a[k] == pytest.approx(b[k], rel=1e-12), (k, a[k], b[k])
        assert a['beam'] == pytest.approx(b['beam'], rel=1e-9) and a['field'] == pytest.approx(b['field'], rel=1e-9)
