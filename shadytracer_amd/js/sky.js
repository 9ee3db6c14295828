'use strict'
/*
 * The sun & sky (the builder's envMode 0; environment.js SUNSKY_FN:110-170, SKY_BAND_TABLE:66-82, sdf.js
 * airMass:96, sunCone in photometry.js:55): sunlight of E0 lux above the air, a 5778 K white, through a plane-parallel
 * atmosphere of air (Rayleigh), haze (Angstrom's law, Henyey-Greenstein g 0.7, single-scattering albedo 0.9) and
 * ozone (absorbing only), in sixteen bands from 380 to 730 nm. The disc is the light left after the sun's air mass;
 * the sky is that light scattered once towards the eye, summed over the column:
 *
 *     L = E0 w P mv (exp(-tau ms) - exp(-tau mv)) / (mv - ms)        (tau mv exp(-tau mv) where mv = ms)
 *
 * Each band's colour is its share of the 5778 K sunlight in linear Rec.709, its XYZ summed over the band at 1 nm.
 * The colour-matching functions here are Wyman, Sloan and Shirley's multi-lobe fit to the CIE 1931 2-degree
 * observer (JCGT 2(2), 2013, eqs. 2-4), not the CIE's table: the bands' colours come within about 3 % of the table's
 * band by band, the sky's and the sun's illuminance within 0.3 % and their chromaticity within 1 %.
 *
 * The device reads the table as literals (csrc/device/pt_sky.h, written by headerText() below); the oracle computes
 * its own in double.
 */

const BANDS = 16, LO = 380, HI = 730
const LUMA = [0.2126, 0.7152, 0.0722]
// IEC 61966-2-1's four-decimal matrix, linear Rec.709 to XYZ, and its inverse
const RGB_TO_XYZ = [[0.4124, 0.3576, 0.1805], LUMA, [0.0193, 0.1192, 0.9505]]
function inverse3(m) {
    const [a, b, c] = m
    const co = [[b[1] * c[2] - b[2] * c[1], a[2] * c[1] - a[1] * c[2], a[1] * b[2] - a[2] * b[1]],
        [b[2] * c[0] - b[0] * c[2], a[0] * c[2] - a[2] * c[0], a[2] * b[0] - a[0] * b[2]],
        [b[0] * c[1] - b[1] * c[0], a[1] * c[0] - a[0] * c[1], a[0] * b[1] - a[1] * b[0]]]
    const det = a[0] * co[0][0] + a[1] * co[1][0] + a[2] * co[2][0]
    return co.map(r => r.map(x => x / det))
}
const XYZ_TO_RGB = inverse3(RGB_TO_XYZ)

// Wyman-Sloan-Shirley's piecewise Gaussian: sigma1 left of the mean, sigma2 right of it
const lobe = (x, mu, s1, s2) => { const t = (x - mu) / (x < mu ? s1 : s2); return Math.exp(-0.5 * t * t) }
function cmf(nm) {
    return [1.056 * lobe(nm, 599.8, 37.9, 31.0) + 0.362 * lobe(nm, 442.0, 16.0, 26.7) - 0.065 * lobe(nm, 501.1, 20.4, 26.2),
        0.821 * lobe(nm, 568.8, 46.9, 40.5) + 0.286 * lobe(nm, 530.9, 16.3, 31.1),
        1.217 * lobe(nm, 437.0, 11.8, 36.0) + 0.681 * lobe(nm, 459.0, 26.0, 13.8)]
}
// Planck's law at 5778 K (arbitrary units), the sunlight above the air
const planck5778 = nm => { const m = nm * 1e-9; return 1 / (m ** 5 * (Math.exp(1.438776877e-2 / (m * 5778)) - 1)) }
// Rayleigh optical depth straight up at sea level, wavelength in micrometres (Hansen & Travis 1974, eq. 2.29)
const rayleighDepth = um => 0.008569 * um ** -4 * (1 + 0.0113 * um ** -2 + 0.00013 * um ** -4)
// the haze's depth relative to 550 nm (Angstrom, exponent 1.3)
const angstrom = nm => (nm / 550) ** -1.3
// ozone: each band's Chappuis-band cross-section (cm^2 a molecule, 223 K; Serdyuchenko et al. 2014, AMT 7, averaged
// over the bands as environment.js OZONE_SIGMA:56 gives them), for 300 Dobson units of 2.6867e16 molecules a cm^2
const OZONE_SIGMA = [6.2648e-24, 3.1169e-23, 1.0195e-22, 2.4886e-22, 5.4883e-22, 1.1517e-21, 2.0211e-21, 3.0497e-21,
    4.2480e-21, 4.5252e-21, 4.7153e-21, 3.4726e-21, 2.3885e-21, 1.5936e-21, 9.8370e-22, 6.3061e-22]
const OZONE_COLUMN = 300 * 2.6867e16

let table = null
/** The sixteen bands: { nm (the middle weighted by light), tauR, haze, ozone, rgb (luminance of all sixteen 1) }. */
function bandTable() {
    if (table) return table
    const step = (HI - LO) / BANDS, raw = []
    for (let k = 0; k < BANDS; k++) {
        const xyz = [0, 0, 0]
        let wl = 0, ws = 0
        for (let l = LO + k * step; l < LO + (k + 1) * step; l += 1) {
            const c = cmf(Math.round(l)), w = planck5778(l)
            for (let i = 0; i < 3; i++) xyz[i] += w * c[i]
            const ww = w * (c[0] + c[1] + c[2])
            wl += ww * l; ws += ww
        }
        raw.push({ nm: wl / ws, rgb: XYZ_TO_RGB.map(r => r[0] * xyz[0] + r[1] * xyz[1] + r[2] * xyz[2]) })
    }
    const tot = raw.reduce((a, b) => a.map((x, i) => x + b.rgb[i]), [0, 0, 0])
    const Y = LUMA[0] * tot[0] + LUMA[1] * tot[1] + LUMA[2] * tot[2]
    table = raw.map((b, k) => ({ nm: b.nm, tauR: rayleighDepth(b.nm / 1000), haze: angstrom(b.nm), ozone: OZONE_SIGMA[k] * OZONE_COLUMN,
        rgb: b.rgb.map(x => x / Y) }))
    return table
}

/** Kasten-Young relative air mass for the cosine of the zenith angle: 1 straight up, finite at the horizon. */
function airMass(cz) {
    const c = Math.min(Math.max(cz, 0), 1)
    const z = Math.acos(c) * 180 / Math.PI
    return 1 / (c + 0.15 * Math.max(93.885 - z, 0.1) ** -1.253)
}

/** The sun's cone for a diameter in degrees, in double: half angle, 1 - cos, sin^2 and solid angle (photometry.js:55). */
function sunCone(diameterDeg) {
    const half = Math.max(diameterDeg, 1e-3) * Math.PI / 360
    const s = Math.sin(half / 2)
    const omc = 2 * s * s
    return { half, oneMinusCos: omc, cos: 1 - omc, sin2: Math.sin(half) ** 2, omega: 2 * Math.PI * omc }
}

/** The reference's sun direction: azimuth from +y towards +x, elevation above the horizon, degrees (sdf.js:72). */
function sunDirection(azDeg, elDeg) {
    const a = azDeg * Math.PI / 180, e = elDeg * Math.PI / 180
    return [Math.cos(e) * Math.sin(a), Math.cos(e) * Math.cos(a), Math.sin(e)]
}

/** The aerosol's depth at 550 nm for a turbidity: 0.05 + 0.1 turbidity. */
const aerosolOf = turbidity => 0.05 + 0.1 * Math.max(turbidity, 0)

/** The disc's radiance, RGB cd/m^2: what the air lets through of E0 over the sun's air mass, spread over pi sin^2. */
function sunRadiance(sun, turbidity, E0, sin2) {
    const T = bandTable(), ms = airMass(Math.max(sun[2], 0)), aer = aerosolOf(turbidity)
    const L = [0, 0, 0]
    for (const b of T) {
        const v = Math.exp(-(b.tauR + aer * b.haze + b.ozone) * ms) / (Math.PI * sin2)
        for (let i = 0; i < 3; i++) L[i] += b.rgb[i] * v
    }
    return L.map(x => Math.max(E0, 0) * x)
}

/** The sky's radiance, RGB cd/m^2, in unit direction d (outside the sun's disc). */
function skyRadiance(d, sun, turbidity, E0) {
    const T = bandTable(), aer = aerosolOf(turbidity)
    const mv = airMass(Math.max(d[2], 0)), ms = airMass(Math.max(sun[2], 0))
    const c = d[0] * sun[0] + d[1] * sun[1] + d[2] * sun[2], g = 0.7
    const pR = 3 / (16 * Math.PI) * (1 + c * c)
    const pM = (1 - g * g) / (4 * Math.PI * (1 + g * g - 2 * g * c) ** 1.5)
    const up = Math.min(Math.max((sun[2] + 0.05) / 0.05, 0), 1)
    const dm = mv - ms
    const L = [0, 0, 0]
    for (const b of T) {
        const ta = aer * b.haze, tau = b.tauR + ta + b.ozone
        const k1 = Math.abs(dm) < 1e-3 ? tau * mv * Math.exp(-tau * mv) : mv * (Math.exp(-tau * ms) - Math.exp(-tau * mv)) / dm
        const v = (b.tauR * pR + 0.9 * ta * pM) / tau * k1 * up
        for (let i = 0; i < 3; i++) L[i] += b.rgb[i] * v
    }
    return L.map(x => Math.max(Math.max(E0, 0) * x, 0))
}

/** The device's table as C++ literals (csrc/device/pt_sky.h). */
function headerText() {
    const T = bandTable(), f = x => x.toPrecision(9) + 'f'
    const row = (name, xs) => `static constexpr float ${name}[${xs.length}] = {\n    ${xs.map(f).join(', ')} };`
    return [row('PT_SKY_TAUR', T.map(b => b.tauR)), row('PT_SKY_HAZE', T.map(b => b.haze)), row('PT_SKY_OZONE', T.map(b => b.ozone)),
        row('PT_SKY_RGB', T.flatMap(b => b.rgb))].join('\n')
}

module.exports = { bandTable, airMass, sunCone, sunDirection, aerosolOf, sunRadiance, skyRadiance, headerText, cmf, BANDS }

if (require.main === module) process.stdout.write(headerText() + '\n')
