"""The optics oracle pinned before it is trusted: published CCTs, Fresnel, a bare Lambertian LED's far field and lux,
flux balance, a plano-convex lens collimating a point source at its front focus. CPU only."""
import math

import numpy as np
import pytest

import oracle
from shadytracer_amd import optics


def test_cct_of_cie_illuminants():
    # CIE illuminant A (x 0.44757, y 0.40745) is 2856 K on the locus; D65 (0.31271, 0.32902) is 6504 K, Duv +0.0032
    T, duv = oracle.cct_duv(0.44757, 0.40745)
    assert abs(T - 2856) < 3 and abs(duv) < 3e-4, (T, duv)
    T, duv = oracle.cct_duv(0.31271, 0.32902)
    assert abs(T - 6504) < 4 and abs(duv - 0.0032) < 3e-4, (T, duv)


def test_cct_duv_round_trip():
    for T in (2700, 3000, 4000, 5000, 6500, 9000):
        for d in (-0.008, 0.0, 0.005):
            t2, d2 = oracle.cct_duv(*optics.cct_duv_to_xy(T, d))
            assert abs(t2 / T - 1) < 2e-4 and abs(d2 - d) < 2e-5, (T, d, t2, d2)


def test_fresnel_unpolarised():
    assert abs(oracle.fresnel_unpolarised(1.0, 1 / 1.5) - 0.04) < 1e-12
    assert oracle.fresnel_unpolarised(0.3, 1.5) == 1.0                     # past the critical angle from inside
    # Brewster's angle: rp = 0, so F = rs^2 / 2
    tb = math.atan(1.5)
    ci, ct = math.cos(tb), math.sqrt(1 - (math.sin(tb) / 1.5) ** 2)
    rs = (ci / 1.5 - ct) / (ci / 1.5 + ct)
    assert abs(oracle.fresnel_unpolarised(ci, 1 / 1.5) - 0.5 * rs * rs) < 1e-12


def rows(far):
    """Pooled intensity per gamma row (cd) from a far field in lm per bin."""
    dG = math.radians(1.0)
    out = []
    for j in range(far.shape[0]):
        lo, hi = max(0.0, (j - 0.5) * dG), min(math.pi, (j + 0.5) * dG)
        out.append(far[j].sum() / (2 * math.pi * (math.cos(lo) - math.cos(hi))))
    return np.array(out)


def test_bare_led_is_lambertian():
    # I(gamma) = (flux / pi) cos(gamma); the lux on the axis at 1 m is I0; nothing is absorbed
    r, F, L = oracle.optics_trace([optics.bare_led(1 << 22)], threads=0, tallies=True)
    r = r[0]
    I = rows(F[0])
    g = np.radians(np.arange(181))
    want = 100 / math.pi * np.maximum(np.cos(g), 0)
    sel = slice(5, 81)
    assert np.max(np.abs(I[sel] / want[sel] - 1)) < 0.03
    assert I[95:].sum() == 0
    assert abs(r['efficiency'] - 1) < 1e-6 and r['absorbed'] < 1e-6 * r['emitted']   # the air lens's rim, 5 km out
    # the field angle hardly moves with I max's noise; the beam angle's 50 % level is set by the brightest (noisiest)
    # row, so it sits a few degrees inside 120
    assert abs(r['field'][0] - 2 * math.degrees(math.acos(0.1))) < 1.5, r['field']
    assert 110 < r['beam'][0] < 121, r['beam']
    assert 1.0 <= r['Imax'] / (100 / math.pi) < 1.15
    centre = L[0, 120:136, 120:136].mean()
    assert abs(centre / (100 / math.pi) - 1) < 0.03, centre
    assert abs(r['cct'] - oracle.cct_duv(0.3457, 0.3585)[0]) < 0.05     # the LED's colour, stored in float


def test_flux_balance_and_reflector_losses():
    sc = optics.bench_scenes(6, log2_min=14, log2_max=15)
    res, _, _ = oracle.optics_trace(sc, threads=0)
    for s, r in zip(sc, res):
        assert abs(r['escaped'] + r['absorbed'] - r['emitted']) < 1e-9 * r['emitted']
        assert r['photons'] == s.photons
        assert 0.5 < r['efficiency'] <= 1.0


def test_plano_convex_lens_collimates_a_point_at_its_front_focus():
    # plano face towards the source: the front focal distance is f - t / n, f = R / (n - 1): the light the lens takes
    # leaves parallel but for spherical aberration -- inside the grid's first degree -- and a source moved off the focus
    # spreads it
    R, n, T, D = 0.02, 1.5, 0.003, 0.008
    f = R / (n - 1)
    at = optics.scene(led_half=1e-6, lens_z=f - T / n, lens_t=T, lens_d=D, c2=-1 / R, ior=n, photons=1 << 19, seed=3)
    off = optics.scene(led_half=1e-6, lens_z=0.8 * (f - T / n), lens_t=T, lens_d=D, c2=-1 / R, ior=n, photons=1 << 19, seed=3)
    r, F, _ = oracle.optics_trace([at, off], threads=0, tallies=True)
    assert r[0]['gammaMax'] == 0 and r[0]['beam'][0] <= 1.0 + 1e-9, r[0]['beam']
    assert r[1]['beam'][0] > 2.0, r[1]['beam']
    # the flux within 1.5 degrees of the axis: a Lambertian source's share inside the cone of rays that enter the plano
    # face at r and leave the glass inside the rim, r + T tan(asin(sin(atan(r / z)) / n)) <= D / 2, less the two
    # faces' Fresnel reflection, (1 - 0.04)^2 at these small angles
    z = f - T / n
    lo, hi = 0.0, D / 2
    for _ in range(60):
        r = (lo + hi) / 2
        if r + T * math.tan(math.asin(math.sin(math.atan(r / z)) / n)) <= D / 2:
            lo = r
        else:
            hi = r
    share = math.sin(math.atan(lo / z)) ** 2 * 100 * (1 - 0.04) ** 2
    near = F[0, :2].sum()
    assert abs(near / share - 1) < 0.02, (near, share)
