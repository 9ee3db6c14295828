'use strict'
/*
 * The far field of an optics trace as luminous intensity, and the beam and field angles read off it (the numbers of
 * lib/pathtracer/optics/farfield.js). The grid is the lighting convention's type C as the trace bins it: nG gamma rows
 * centred on 0, 180 / (nG - 1), ..., 180 degrees, the end ones half wide, nC C columns centred on 0, 360 / nC, ....
 * A bin's solid angle is (2 pi / nC) (cos lo - cos hi); a row's columns are pooled at the poles, and for a round beam
 * ('rotational') everywhere. A width is where the C plane through C and C + 180 first falls below the level walking out
 * from its maximum, linear between the two bins it falls between.
 */
const RAD = Math.PI / 180

function omegaOf(nC, nG) {
    const dG = 180 / (nG - 1), w = new Float64Array(nG)
    for (let j = 0; j < nG; j++) {
        const lo = Math.max(0, (j - 0.5) * dG) * RAD, hi = Math.min(180, (j + 0.5) * dG) * RAD
        w[j] = (2 * Math.PI / nC) * (Math.cos(lo) - Math.cos(hi))
    }
    return w
}

/** cd per bin (gamma-major, like the grid) from a far field { nC, nG, flux }. */
function intensity(far, { symmetry = 'none' } = {}) {
    const { nC, nG, flux } = far
    const w = omegaOf(nC, nG), I = new Float64Array(nC * nG)
    for (let j = 0; j < nG; j++) {
        const pool = j === 0 || j === nG - 1 || symmetry === 'rotational'
        let f = 0
        for (let i = 0; i < nC; i++) f += flux[j * nC + i]
        for (let i = 0; i < nC; i++) I[j * nC + i] = pool ? f / (w[j] * nC) : flux[j * nC + i] / w[j]
    }
    return { nC, nG, I, dC: 360 / nC, dG: 180 / (nG - 1), symmetry }
}

/** The C plane through C and C + 180: { a: signed angles -180..180, I } (C + 180 on the minus side). */
function plane(R, C) {
    const at = (j, c) => {
        const x = ((c % 360) + 360) % 360 / R.dC, i0 = Math.floor(x) % R.nC, t = x - Math.floor(x)
        return R.I[j * R.nC + i0] * (1 - t) + R.I[j * R.nC + (i0 + 1) % R.nC] * t
    }
    const a = [], I = []
    for (let j = R.nG - 1; j >= 1; j--) { a.push(-j * R.dG); I.push(at(j, C + 180)) }
    for (let j = 0; j < R.nG; j++) { a.push(j * R.dG); I.push(at(j, C)) }
    return { a, I }
}

/** The full width (degrees) of plane C at `level` of its maximum: 0.5 the beam angle, 0.1 the field angle. */
function width(R, C = 0, level = 0.5) {
    const { a, I } = plane(R, C)
    let p = 0
    for (let k = 1; k < I.length; k++) if (I[k] > I[p]) p = k
    const cross = step => {
        for (let k = p; k + step >= 0 && k + step < I.length; k += step) {
            if (I[k + step] < level * I[p]) {
                const t = I[k] === I[k + step] ? 0 : (I[k] - level * I[p]) / (I[k] - I[k + step])
                return a[k] + (a[k + step] - a[k]) * t
            }
        }
        return null
    }
    const lo = cross(-1), hi = cross(1)
    return lo === null || hi === null ? null : hi - lo
}

/** The grid as CSV text: a comment line, a header of gamma and one column per C (each followed by its standard error
 *  when R.se is given), then one row per gamma (optics/farfield.js intensityCSV :170). */
function intensityCSV(R, unit = 'cd') {
    const fmt = x => String(+x.toFixed(4))
    const head = ['gamma']
    for (let i = 0; i < R.nC; i++) { head.push('C' + fmt(i * R.dC)); if (R.se) head.push('se_C' + fmt(i * R.dC)) }
    const rows = [`# type C grid, ${unit}: a row per gamma (degrees from the photometric axis), a column per C (degrees about it from C0)`,
        head.join(',')]
    for (let j = 0; j < R.nG; j++) {
        const r = [fmt(j * R.dG)]
        for (let i = 0; i < R.nC; i++) { r.push(R.I[j * R.nC + i]); if (R.se) r.push(R.se[j * R.nC + i]) }
        rows.push(r.join(','))
    }
    return rows.join('\n')
}

module.exports = { intensity, plane, width, omegaOf, intensityCSV }
