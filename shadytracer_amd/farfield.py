"""The far field of an optics trace as luminous intensity, and the numbers read off it (optics/farfield.js).

The grid is the lighting convention's type C, as the trace bins it: nG gamma rows centred on 0, 180 / (nG - 1), ...,
180 degrees from the photometric axis, the first and last half as wide (caps at the poles), and nC C columns about the
axis centred on 0, 360 / nC, .... A bin's solid angle is (2 pi / nC) (cos lo - cos hi), lo and hi its gamma edges
(farfield.js farGeometry :31-40). Intensity is flux over solid angle; at the poles, and for a round beam
('rotational'), a row's columns are pooled: the row's flux over the row's solid angle (farfield.js intensity :64-83).
The beam and field angles are the full widths at 50 % and 10 % of the largest intensity, each side found walking out
from the maximum to where the curve first falls below the level, linear between the two bins it falls between
(farfield.js crossing :104-115, farStats).
"""
import math

import numpy as np


def geometry(nC, nG):
    """Bin widths in degrees and each gamma row's solid angle per C bin (sr)."""
    dC, dG = 360.0 / nC, 180.0 / (nG - 1)
    j = np.arange(nG)
    lo = np.radians(np.maximum(0.0, (j - 0.5) * dG))
    hi = np.radians(np.minimum(180.0, (j + 0.5) * dG))
    return dC, dG, (2.0 * math.pi / nC) * (np.cos(lo) - np.cos(hi))


def intensity(far, symmetry='none'):
    """Candela per (gamma, C) bin, [nG, nC], from a far field {'flux': [nG, nC] lm, 'nC', 'nG'}."""
    F = np.asarray(far['flux'], float)
    nG, nC = F.shape
    _, _, omega = geometry(nC, nG)
    I = F / omega[:, None]
    pooled = F.sum(axis=1) / (omega * nC)
    rows = range(nG) if symmetry == 'rotational' else (0, nG - 1)
    for j in rows:
        I[j, :] = pooled[j]
    return I


def plane(I, C):
    """The C plane through C and C + 180 as (angles, intensities): gamma -180..180 signed, C + 180 on the minus side,
    a C between columns interpolated linearly."""
    nG, nC = I.shape
    dC, dG = 360.0 / nC, 180.0 / (nG - 1)

    def at(j, c):
        x = (c % 360.0) / dC
        i0 = int(math.floor(x)) % nC
        t = x - math.floor(x)
        return I[j, i0] * (1.0 - t) + I[j, (i0 + 1) % nC] * t
    a = [-j * dG for j in range(nG - 1, 0, -1)] + [j * dG for j in range(nG)]
    v = [at(j, C + 180.0) for j in range(nG - 1, 0, -1)] + [at(j, C) for j in range(nG)]
    return np.array(a), np.array(v)


def _crossing(a, v, p, step, level):
    k = p
    while 0 <= k + step < len(v):
        if v[k + step] < level:
            t = 0.0 if v[k] == v[k + step] else (v[k] - level) / (v[k] - v[k + step])
            return a[k] + (a[k + step] - a[k]) * t
        k += step
    return None


def width(I, C=0.0, level=0.5):
    """The full width (degrees) of plane C where it falls to `level` of its maximum; None where it never does."""
    a, v = plane(I, C)
    p = int(np.argmax(v))
    lo, hi = _crossing(a, v, p, -1, level * v[p]), _crossing(a, v, p, 1, level * v[p])
    return None if lo is None or hi is None else hi - lo


def i0_fit(I, max_gamma=20.0):
    """I0 of a fit I = I0 cos(gamma) over the rows within max_gamma of the axis (pooled rows), least squares."""
    nG = I.shape[0]
    dG = 180.0 / (nG - 1)
    g = np.radians(np.arange(nG) * dG)
    rows = np.arange(nG)[np.arange(nG) * dG <= max_gamma + 1e-9]
    Ir = I[rows].mean(axis=1)
    c = np.cos(g[rows])
    return float((Ir * c).sum() / (c * c).sum())


def intensity_csv(I, unit='cd', se=None):
    """The grid as CSV text (farfield.js intensityCSV :170): a comment line, a header of gamma and one column per C
    (each followed by its standard error when `se` is given), then one row per gamma."""
    nG, nC = I.shape
    dC, dG = 360.0 / nC, 180.0 / (nG - 1)
    fmt = lambda x: ('%.4f' % x).rstrip('0').rstrip('.')
    head = ['gamma']
    for i in range(nC):
        head.append('C' + fmt(i * dC))
        if se is not None:
            head.append('se_C' + fmt(i * dC))
    rows = ['# type C grid, %s: a row per gamma (degrees from the photometric axis), a column per C (degrees about it from '
            'C0)' % unit, ','.join(head)]
    for j in range(nG):
        r = [fmt(j * dG)]
        for i in range(nC):
            r.append(repr(float(I[j, i])))
            if se is not None:
                r.append(repr(float(se[j, i])))
        rows.append(','.join(r))
    return '\n'.join(rows)
