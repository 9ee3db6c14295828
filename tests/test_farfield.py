"""The far field's numbers (shadytracer_amd/farfield.py and js/farfield.js; optics/farfield.js) on exactly known far
fields: a source's flux binned on the type C grid in closed form, read back as intensity and beam / field angles."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

from shadytracer_amd import farfield

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def grid(share, nC=24, nG=37, flux=1000.0):
    dG = 180.0 / (nG - 1)
    F = np.zeros((nG, nC))
    for j in range(nG):
        lo, hi = math.radians(max(0.0, (j - 0.5) * dG)), math.radians(min(180.0, (j + 0.5) * dG))
        F[j, :] = flux * (share(hi) - share(lo)) / nC
    return {'flux': F, 'nC': nC, 'nG': nG}


def cos_n(n):
    return lambda g: 1.0 - max(math.cos(min(g, math.pi / 2)), 0.0) ** (n + 1)


def test_solid_angles_cover_the_sphere():
    for nC, nG in ((24, 37), (72, 181), (1, 2)):
        _, _, w = farfield.geometry(nC, nG)
        assert w.sum() * nC == pytest.approx(4 * math.pi, rel=1e-12)


def test_lambertian_on_a_fine_grid_is_the_closed_form():
    # I = (flux / pi) cos(gamma): the half level at 60 degrees, the tenth at acos 0.1; I0 = flux / pi
    I = farfield.intensity(grid(cos_n(1), nC=72, nG=721), 'rotational')
    assert farfield.width(I, 0.0, 0.5) == pytest.approx(120.0, abs=0.02)
    assert farfield.width(I, 0.0, 0.1) == pytest.approx(2 * math.degrees(math.acos(0.1)), abs=0.05)
    assert farfield.i0_fit(I) == pytest.approx(1000 / math.pi, rel=1e-4)


@pytest.mark.parametrize('n', [1, 5, 20])
def test_cos_n_beam_converges_with_the_grid(n):
    closed = 2 * math.degrees(math.acos(0.5 ** (1 / n)))
    coarse = farfield.width(farfield.intensity(grid(cos_n(n)), 'rotational'))
    fine = farfield.width(farfield.intensity(grid(cos_n(n), nG=1441), 'rotational'))
    assert fine == pytest.approx(closed, abs=0.02)
    assert coarse == pytest.approx(closed, rel=0.02)


def test_a_tilted_beam_reads_in_its_plane():
    # all the flux in one (gamma, C) bin, 30 degrees off the axis at C 90: the C90 plane peaks there, the C0 plane does
    # not see it
    F = np.zeros((37, 24))
    F[6, 6] = 1000.0
    I = farfield.intensity({'flux': F, 'nC': 24, 'nG': 37})
    a, v = farfield.plane(I, 90.0)
    assert a[int(np.argmax(v))] == pytest.approx(30.0)
    _, v0 = farfield.plane(I, 0.0)
    assert v0.max() == 0.0


def test_javascript_reads_the_same_numbers():
    far = grid(cos_n(5))
    src = """
const f = require(%s)
const far = %s
const I = f.intensity({ nC: far.nC, nG: far.nG, flux: Float64Array.from(far.flux) }, { symmetry: 'rotational' })
process.stdout.write(JSON.stringify([f.width(I, 0, 0.5), f.width(I, 0, 0.1)]))
""" % (json.dumps(os.path.join(ROOT, 'shadytracer_amd', 'js', 'farfield.js')),
       json.dumps({'nC': 24, 'nG': 37, 'flux': far['flux'].ravel().tolist()}))
    out = json.loads(subprocess.run(['node', '-e', src], check=True, stdout=subprocess.PIPE, text=True).stdout)
    I = farfield.intensity(far, 'rotational')
    assert out[0] == pytest.approx(farfield.width(I, 0.0, 0.5), rel=1e-12)
    assert out[1] == pytest.approx(farfield.width(I, 0.0, 0.1), rel=1e-12)


def test_intensity_csv_holds_the_grid():
    I = farfield.intensity(grid(cos_n(1)), 'rotational')
    text = farfield.intensity_csv(I)
    lines = text.split('\n')
    assert lines[0].startswith('#') and lines[1].split(',')[:3] == ['gamma', 'C0', 'C15']
    back = np.array([[float(x) for x in l.split(',')] for l in lines[2:]])
    assert back.shape == (37, 25) and np.allclose(back[:, 0], np.arange(37) * 5.0)
    assert np.array_equal(back[:, 1:], I)


def test_detector_map_of_uniform_hits():
    # 2e5 hits of equal flux uniform over a 1 x 0.5 m rectangle on the floor: 2000 lx everywhere, the floor's frame
    from shadytracer_amd import detmap
    rng = np.random.default_rng(1)
    n = 200000
    H = np.zeros((n, 8))
    H[:, 0], H[:, 1], H[:, 3] = rng.uniform(-0.5, 0.5, n), rng.uniform(-0.25, 0.25, n), 1000.0 / n
    m = detmap.detector_map(H, cells=20, rng={'a': (-0.5, 0.5), 'b': (-0.25, 0.25)})
    assert (m['nx'], m['ny']) == (20, 10) and np.allclose(m['normal'], [0, 0, 1])
    # the range is about the hits' centre (their mean, a hair off the origin): a few hits at the edge fall outside
    assert m['total'] == pytest.approx(1000.0, rel=2e-3) and m['E'].sum() * m['cell'] ** 2 == pytest.approx(m['total'])
    assert m['meanLit'] == pytest.approx(2000.0, rel=2e-3)
    # each cell holds ~1000 hits: 3 % noise
    assert m['uniformity'] > 0.85 and m['max'] / m['meanLit'] < 1.15
    # a tilted detector: the plane's frame follows the hits
    t = np.radians(30)
    R = np.array([[1, 0, 0], [0, np.cos(t), -np.sin(t)], [0, np.sin(t), np.cos(t)]])
    H2 = H.copy()
    H2[:, :3] = H[:, :3] @ R.T
    m2 = detmap.detector_map(H2, cells=20)
    assert abs(abs(m2['normal'] @ R[:, 2]) - 1) < 1e-6
    assert m2['total'] == pytest.approx(1000.0)


def test_javascript_detector_map_agrees(tmp_path):
    from shadytracer_amd import detmap
    rng = np.random.default_rng(2)
    n = 5000
    H = np.zeros((n, 8))
    H[:, 0], H[:, 1], H[:, 2], H[:, 3] = rng.normal(0, 0.2, n), rng.normal(0, 0.1, n), 1.0, 1.0 / n
    js = os.path.join(ROOT, 'shadytracer_amd', 'js', 'detmap.js')
    path = tmp_path / 'hits.json'
    path.write_text(json.dumps(H.tolist()))
    src = "const {detectorMap} = require(%s); const H = JSON.parse(require('fs').readFileSync(%s, 'utf8'));\n" % (
        json.dumps(js), json.dumps(str(path))) + \
          "const m = detectorMap(H.map(h => ({pos: h.slice(0, 3), flux: h[3]})), {cells: 16});\n" + \
          "process.stdout.write(JSON.stringify({E: Array.from(m.E), nx: m.nx, ny: m.ny, normal: m.normal, total: m.total}))"
    out = json.loads(subprocess.run(['node', '-e', src], check=True, stdout=subprocess.PIPE, text=True).stdout)
    m = detmap.detector_map(H, cells=16)
    assert (out['nx'], out['ny']) == (m['nx'], m['ny'])
    assert np.allclose(out['normal'], m['normal']) and out['total'] == pytest.approx(m['total'])
    assert np.allclose(np.array(out['E']).reshape(m['ny'], m['nx']), m['E'], rtol=1e-9)
