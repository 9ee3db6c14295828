'use strict'
// The emitter's colour, camera and refusal rules without a GPU. Prints JSON for tests/test_emitter.py:
//   locus       Krystek's fit of the Planckian locus in xy at the tabulated blackbody temperatures
//   duvStep     the 1960-uv distance of a Duv of 0.01 from the locus at 4000 K, and whether it is above (greener)
//   d65, blue   the RGB of D65's chromaticity and of a royal blue outside Rec.709 (luminance, XYZ round trip)
//   camera      the orbit view turned into the engine's pinhole
//   refusals    what each document is refused for, or null
const path = require('path')
const js = path.join(__dirname, '..', '..', 'shadytracer_amd', 'js')
const H = require(path.join(js, 'hipgen'))

const out = {}
const xyOfUv = (u, v) => { const d = 2 * u - 8 * v + 4; return [3 * u / d, 2 * v / d] }
const uvOfXy = (x, y) => { const d = -2 * x + 12 * y + 3; return [4 * x / d, 6 * y / d] }
out.locus = [2000, 2856, 3000, 4000, 5000, 6500, 10000].map(T => [T, ...H.cctDuvToXy(T, 0)])
const p0 = uvOfXy(...H.cctDuvToXy(4000, 0)), p1 = uvOfXy(...H.cctDuvToXy(4000, 0.01))
out.duvStep = [Math.hypot(p1[0] - p0[0], p1[1] - p0[1]), p1[1] > p0[1]]
out.d65 = H.lightRGB({ colourBy: 'xy', xy: { x: 0.3127, y: 0.3290 } })
out.blue = H.lightRGB({ colourBy: 'xy', xy: { x: 0.15, y: 0.025 } })
out.cct = H.lightRGB({ colourBy: 'cct', cct: 2856 })
out.rgb = H.lightRGB({ colour: { r: 2, g: 1, b: 0 } })
void xyOfUv

// a camera 450 mm from a target 120 mm up, 60 degrees round from +x, 75 degrees down from +z, 40 degrees tall
const view = { unit: 'mm', view: { camera: { mode: 'orbit', camtype: 0, fov: 40, target: [0, 0, 120], azimuth: 60, elevation: 75, distance: 450, up: [0, 0, 1], aperture: 0 } } }
out.camera = H.cameraOf(view, 400, 200, () => {})

const doc = (extra) => Object.assign({ format: 'raytracer-scene', version: 31, unit: 'm', blocks: [{ id: 'b', builtin: 'sphere@1', params: { radius: 0.5 } }],
    materials: [{ id: 'm', params: { baseColor: { r: 0.5, g: 0.5, b: 0.5 } } }],
    root: { id: 'root', kind: 'group', children: [{ id: 'n', kind: 'object', type: 'code', block: 'b', material: 'm', params: {} }] },
    environment: { envMode: 1, studioIlluminance: 100 } }, extra)
const refusal = d => { try { H.generate(d); return null } catch (e) { return String(e.message) } }
const mat = p => [{ id: 'm', params: Object.assign({ baseColor: { r: 0.5, g: 0.5, b: 0.5 } }, p.params || {}), code: p.code, relief: p.relief }]
const light = L => ({ id: 'root', kind: 'group', children: [{ id: 'L', kind: 'object', type: 'light', params: { width: 1, height: 1 }, light: L }] })
out.refusals = {
    studio: refusal(doc({})),
    sunsky: refusal(doc({ environment: { envMode: 0 } })),
    nosetting: refusal(doc({ environment: undefined })),
    envcode: refusal(doc({ environment: { envMode: 2 } })),
    comments: refusal(doc({ materials: mat({ code: '// only a comment\n   // and another\n' }) })),
    code: refusal(doc({ materials: mat({ code: '// a texture\nm.baseColor = myTexture(p);\n' }) })),
    translated: refusal(doc({ materials: mat({ code: '// a checker\nm.baseColor = VEC3(step(0.0, sin(p.x)));\n' }) })),
    coating: refusal(doc({ materials: mat({ params: { coating: 1 } }) })),
    relief: refusal(doc({ materials: mat({ relief: { code: 'h = heightAt(0u, q.xy).x;', params: { amplitude: 1 } } }) })),
    collimated: refusal(doc({ root: light({ shape: 'rectangle', distribution: 'collimated', flux: 10 }) })),
    cone: refusal(doc({ root: light({ shape: 'rectangle', halfAngle: 30 }) })),
    lambertian: refusal(doc({ root: light({ shape: 'rectangle', colourBy: 'cct', cct: 3000, flux: 100 }) })),
    unusedcoating: refusal(doc({ materials: mat({ params: { coating: 1 } }).concat([{ id: 'spare', params: {} }]),
        root: { id: 'root', kind: 'group', children: [{ id: 'n', kind: 'object', type: 'code', block: 'b', params: {} }] } })),
}
const approx = H.generate(doc({ environment: { envMode: 0 }, materials: mat({ params: { anisotropy: 0.8 } }) }), { approximate: true })
out.approximations = approx.approximations
// a light's flux: pi L A on its one face (1 m^2 here)
const lit = H.generate(doc({ root: light({ shape: 'rectangle', colourBy: 'cct', cct: 3000, flux: 100 }) }))
out.lightLe = lit.faces[0].le
// the sun & sky at its defaults but for the sun's place, and at three elevations; the sky's band table and the
// device's copy of it
const S = require(path.join(js, 'sky'))
out.sunsky = [45, 15, 5, -3.44].map(el => H.generate(doc({ environment: { envMode: 0, uSunAzimuth: 30, uSunElevation: el } })).env)
out.skyBands = S.bandTable()
out.skyHeader = S.headerText()
out.skyJs = [[0, 0, 1], [0.6, 0, 0.8], [0, -0.96, 0.28], [0.3, 0.3, -0.905]].map(d => d.map(x => x / Math.hypot(...d)))
    .map(d => S.skyRadiance(d, out.sunsky[0].sunDir, 0.7, 128000))
out.studioEnv = H.generate(doc({ environment: { envMode: 1, studioIlluminance: 300, studioTop: 0.6, studioColor: { r: 1, g: 0.5, b: 0.25 } } })).env
process.stdout.write(JSON.stringify(out))
