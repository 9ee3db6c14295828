'use strict'
// The darkroom's host mirror (shadytracer_amd/js/darkroom.js) evaluated on the cases tests/test_darkroom.py checks;
// prints one JSON object.
const D = require('../../shadytracer_amd/js/darkroom.js')

const apply = (m, v) => m.map(r => r[0] * v[0] + r[1] * v[1] + r[2] * v[2])
const out = {}
out.wbDefault = D.WB_DEFAULT
out.d65 = D.D65
out.illuminantA = D.cctDuvToXy(2856, 0)
out.white = {}
for (const [T, tint] of [[3000, 0], [2700, 5], [9000, -4]]) {
    const xy = D.whiteXy(T, tint)
    const rgb = D.xyzToRgb(D.xyToXYZ(xy[0], xy[1], 1))
    out.white[`${T},${tint}`] = apply(D.whiteBalanceMatrix(T, tint), rgb)
}
out.moved = D.chromaticity(D.rgbToXYZ(apply(D.whiteBalanceMatrix(3000), D.xyzToRgb(D.xyToXYZ(0.40, 0.45, 1)))))
out.gained = apply(D.whiteBalanceMatrix(D.WB_DEFAULT.temp, 0, [2, 1, 0.5]), [1, 1, 1])
out.plain = apply(D.whiteBalanceMatrix(D.WB_DEFAULT.temp, 0), [1, 1, 1])
const lamp = D.xyzToRgb(D.xyToXYZ(...D.cctDuvToXy(3000, 0), 1)).map(c => c * 40)
out.eyedropper = D.whiteFromRGB(lamp)
out.balancedWall = apply(D.whiteBalanceMatrix(out.eyedropper.temp, out.eyedropper.tint), lamp)
out.refused = [D.whiteFromRGB([0, 0, 0]), D.whiteFromRGB([1, -0.1, 1]), D.whiteFromRGB([0, 0, 1])].map(r => r.error || null)
const bits = s => D.lightUniforms({ ...D.LIGHT_DEFAULTS, ...s }).lightOn
const bend = { m: [[0, 0], [0.5, 0.6], [1, 1]] }
out.bits = {
    defaults: bits({}), empty: D.lightUniforms({}).lightOn, wbTemp: bits({ wbTemp: 3000 }), wbTint: bits({ wbTint: 1 }),
    gainB: bits({ gainB: 1.1 }), eq3: bits({ eq3: 0.5 }), labVib: bits({ labVib: 0.2 }), curve7: bits({ toneMap: 7 }),
    otherCurves: [0, 1, 2, 3, 4, 5, 6].map(t => bits({ toneMap: t })), levels: bits({ lvGamma: 1.5 }), curves: bits({ curves: bend }),
    curvesLuma: bits({ curves: bend, curveLuma: true }), lumaAlone: bits({ curveLuma: true }), cdl: bits({ gainY: 0.2 }),
    vignette: bits({ vigAmount: -0.5 }), grain: bits({ grainAmount: 0.3 }), dither: bits({ dither: true }), transfer: bits({ transfer: 1 }),
}
const u = D.lightUniforms({ ...D.LIGHT_DEFAULTS, wbTemp: 3000, gainR: 1.2, eq0: 1, eq7: -1, labHue: 90, lvGamma: 2 })
out.uniforms = { wb: u.wb, wantWb: [].concat(...D.whiteBalanceMatrix(3000, 0, [1.2, 1, 1])), eq: u.eq, lab: u.lab,
    tone7: D.lightUniforms(D.LIGHT_DEFAULTS).tone7, levels: u.levels }
// the curves
const points = [[0, 0.1], [0.2, 0.15], [0.25, 0.6], [0.6, 0.62], [0.8, 0.62], [1, 0.95]]
const f = D.curveFunction(points)
out.curve = { points, at: points.map(p => f(p[0])), xs: [], ys: [] }
for (let i = 0; i <= 2000; i++) { out.curve.xs.push(i / 2000); out.curve.ys.push(f(i / 2000)) }
out.straightLut = Array.from(D.bakeLut({}))
out.bentLut = Array.from(D.bakeLut(bend))
out.normalized = D.normalizeCurve([[0.5, 2], [-1, 0.2], [0.501, 0.7], 'x', [1, 1]])
const rest = D.cdlFromWheels({ liftX: 0, liftY: 0, liftM: 0, gammaX: 0, gammaY: 0, gammaM: 0, gainX: 0, gainY: 0, gainM: 0 })
const lift = D.cdlFromWheels({ liftX: 0, liftY: 0, liftM: 0.5, gammaX: 0, gammaY: 0, gammaM: 0, gainX: 0, gainY: 0, gainM: 1 })
out.cdl = { rest, lift, tint: D.wheelTint(0.6, -0.3), red: D.wheelTint(1, 0) }
out.step = [D.histogramStep(1920, 1080), D.histogramStep(7680, 4320), D.histogramStep(100, 100)]
// recipes
const look = { ...D.DEVELOP_DEFAULTS(), wbTemp: 3400, labVib: 0.3, dither: true, curves: { m: [[0.5, 0.6], [0, 0], [1, 1]] } }
const text = JSON.stringify(D.toRecipe(look))
const back = D.recipeFromText(text)
const odd = D.recipeFromText(JSON.stringify({ settings: { wbTemp: '3200', labVib: true, dither: 1, mystery: 4, curves: 'x' } }))
const { encodePNG } = require('../../shadytracer_amd/js/png.js')
const png = encodePNG(new Uint8Array(4 * 4 * 4), 4, 4, { text: [['Software', 's'], ['Comment', JSON.stringify({ format: 'png', ev: 9.25, settings: look })]] })
const out2 = { recipe: JSON.parse(text), back, odd, fromPicture: D.recipeFromPicture(png), notRecipe: D.recipeFromText('{"kind":"scene"}'),
    notJson: D.recipeFromText('{'), bare: D.recipeFromPicture(encodePNG(new Uint8Array(16), 2, 2)) }
out.recipes = out2
console.log(JSON.stringify(out))
