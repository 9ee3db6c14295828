'use strict'
/*
 * createPathTracer -- the facade of lib/pathtracer (lib/pathtracer/index.js:90; README.md:38-124) over the MI355X
 * engine, for the calls a host makes to render and measure:
 *
 *   const RT = createPathTracer({ gpu, scene, CAMERA, width, height })   gpu: { device } (no WebGPU device here)
 *   await RT.init()
 *   await RT.step(n)                 n samples; resolves when the GPU is done (index.js:1005-1007)
 *   RT.set(values, reset) -> unknownKeys (the engine's uniforms and the reference's lighting settings); RT.get(name)
 *   await RT.setScene(scene, { keepPicture }) -> ms; RT.updateScene(partial) -> refusedKeys; RT.invalidate()
 *   RT.imageStats(); RT.measure(x0, y0, x1, y1); RT.readRegion(...); RT.size(); RT.setSize(w, h)
 *
 * CAMERA: { pos, target, fov (degrees, vertical), up (default +z) } or the engine's { pos, right, up, fwd, fx, fy }.
 * The picture's rows count from the bottom, as the engine's pixel index does (genrays.js:98-99).
 */
const { RAYTRACER, KNOWN } = require('./raytrace_main')
const { camera } = require('./configs')
const { environmentOf, UNIT_METRES } = require('./hipgen')

function cameraUniforms(C, w, h) {
    if (!C) return null
    if (C.right && C.fwd) return Object.assign({ type: 0 }, C, { fx: C.fy * w / h })
    return camera(C.pos, C.target || [0, 0, 0], C.fov || 60, w, h)
}

// The reference's lighting settings (core/uniforms.schema.js Lighting :420-442 and the test environment :405-413), which
// steer the environment of a scene that has none of its own: applied through the emitter's environmentOf onto the
// engine's PtEnv, over what the scene's document said.
const LIGHTING = new Set(['envMode', 'studioIlluminance', 'studioColor', 'studioTop', 'sunHint', 'sunDiameter', 'uSunAzimuth',
    'uSunElevation', 'sunOutside', 'uSkyTurbidity', 'envSun', 'envSky', 'envUniform'])
// what updateScene changes without a compile (lib/pathtracer/README.md updateScene): here the radius, the unit and the
// haze box's values (a box of the emitter's form: centre, half, sigmaS, g, albedo, metres)
const SCENE_FIELDS = new Set(['envRadius', 'unit', 'haze'])

function createPathTracer(opts = {}) {
    const width = opts.width || 256, height = opts.height || 256
    const R = RAYTRACER(opts.gpu || {}, { width, height }, opts.Q, null)
    let CAM = opts.CAMERA || null
    let reset = true
    let lighting = {}               // the lighting settings set so far, over the scene's own environment
    const RT = {}
    RT.init = async () => {
        await R.init()
        if (opts.scene) await RT.setScene(opts.scene)
        return RT
    }
    function applyLighting() {
        const sc = R.scene()
        if (!sc || !Object.keys(lighting).length) return
        const base = Object.assign({}, sc.envSettings || {})
        const E = Object.assign(base, lighting)
        // a scene written with the engine's own sky: the reference's settings take its place once any is set
        if (base.sky !== undefined) delete E.sky
        const env = E.envUniform > 0 ? environmentOf({ sky: 'sphere', uniformSky: E.envUniform }, sc.envDoc || { unit: 'm' })
            : environmentOf(E, sc.envDoc || { unit: 'm' })
        R.uniforms().env = env
    }
    // Resolves with the compile's milliseconds; {keepPicture} keeps the samples so far (index.js:757)
    RT.setScene = async (scene, { keepPicture = false } = {}) => {
        const ms = await R.setScene(scene, { keepPicture })
        applyLighting()
        if (!keepPicture) reset = true
        return ms
    }
    // Unknown names are returned, not thrown (index.js set: a typo must not be a silent no-op, nor an exception mid-frame)
    RT.set = (values, doReset = true) => {
        const unknown = []
        const U = R.uniforms()
        let lit = false
        for (const [k, v] of Object.entries(values || {})) {
            if (k === 'CAMERA' || k === 'camera') { CAM = v; continue }
            if (LIGHTING.has(k)) { lighting[k] = v; lit = true; continue }
            if (!KNOWN.has(k)) { unknown.push(k); continue }
            U[k] = v
        }
        if (lit) applyLighting()
        if (doReset) reset = true
        return unknown
    }
    RT.get = (k) => LIGHTING.has(k) ? lighting[k] : R.uniforms()[k]
    // what the scene says about itself that is not code, changed without a compile; returns the keys refused
    RT.updateScene = (partial = {}) => {
        const refused = []
        const U = R.uniforms(), sc = R.scene()
        for (const [k, v] of Object.entries(partial)) {
            if (!SCENE_FIELDS.has(k) || !sc) { refused.push(k); continue }
            if (k === 'envRadius') { const m = (sc.unit || 1) * (v === null ? 400 / (sc.unit || 1) : v); U.envRadius = m; sc.envRadius = m }
            else if (k === 'unit') { if (!UNIT_METRES[v]) { refused.push(k); continue } sc.unit = UNIT_METRES[v] }
            else if (k === 'haze') { U.haze = v || {}; sc.haze = v }
        }
        reset = true
        return refused
    }
    // the scene changed: what has accumulated is of another picture
    RT.invalidate = () => { reset = true }
    RT.size = () => ({ width: R.uniforms().width, height: R.uniforms().height })
    RT.setSize = (w, h) => { R.ensureBuffers(w, h); reset = true }
    RT.step = async (n = 1) => {
        const U = R.uniforms()
        const cam = cameraUniforms(CAM, U.width, U.height)
        if (cam) U.cam = cam
        if (reset) { await R.trace(null, U, 0); R.clearAccums(); reset = false }
        await R.trace(null, U, n)
    }
    RT.imageStats = () => R.imageStats()
    RT.readRegion = (...a) => R.readRegion(...a)
    // the render's light as a file, 'pfm' | 'hdr' | 'exr' (darkroom/export/index.js: the light formats): a Buffer
    RT.saveLight = (format, opts) => R.saveLight(format, opts)
    // the developed picture (exposure, tone curve, sRGB) and as a PNG file
    RT.develop = opts => R.develop(opts)
    RT.develop16 = opts => R.develop16(opts)
    RT.developHistogram = (opts, step) => R.developHistogram(opts, step)
    RT.developView = (opts, out) => R.developView(opts, out)
    RT.readAccum = () => R.readAccum()
    RT.savePicture = (format, opts) => R.savePicture(format, opts)
    RT.readPixel = (...a) => R.readPixel(...a)
    RT.measure = (x0, y0, x1, y1) => R.readRegion(x0, y0, x1, y1)
    RT.wavefrontLog = () => R.wavefrontLog()
    // for createOptics(RT): the trace over this scene, and the scene it runs on
    RT.opticsRun = (...a) => R.opticsRun(...a)
    RT.scene = () => R.scene()
    RT.dispose = () => R.releaseBuffers()
    return RT
}

module.exports = { createPathTracer, RAYTRACER }
