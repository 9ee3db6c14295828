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
