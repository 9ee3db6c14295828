'use strict'
/*
 * createPathTracer -- the facade of lib/pathtracer (lib/pathtracer/index.js:90; README.md:38-124) over the MI355X
 * engine, for the calls a host makes to render and measure:
 *
 *   const RT = createPathTracer({ gpu, scene, CAMERA, width, height })   gpu: { device } (no WebGPU device here)
 *   await RT.init()
 *   await RT.step(n)                 n samples; resolves when the GPU is done (index.js:1005-1007)
 *   RT.set(values, reset) -> unknownKeys; RT.get(name)
 *   RT.setScene(scene) -> ms; RT.imageStats(); RT.measure(x0, y0, x1, y1); RT.readRegion(...); RT.size(); RT.setSize(w, h)
 *
 * CAMERA: { pos, target, fov (degrees, vertical), up (default +z) } or the engine's { pos, right, up, fwd, fx, fy }.
 * The picture's rows count from the bottom, as the engine's pixel index does (genrays.js:98-99).
 */
const { RAYTRACER, KNOWN } = require('./raytrace_main')
const { camera } = require('./configs')

function cameraUniforms(C, w, h) {
    if (!C) return null
    if (C.right && C.fwd) return Object.assign({ type: 0 }, C, { fx: C.fy * w / h })
    return camera(C.pos, C.target || [0, 0, 0], C.fov || 60, w, h)
}

function createPathTracer(opts = {}) {
    const width = opts.width || 256, height = opts.height || 256
    const R = RAYTRACER(opts.gpu || {}, { width, height }, opts.Q, null)
    let CAM = opts.CAMERA || null
    let reset = true
    const RT = {}
    RT.init = async () => {
        await R.init()
        if (opts.scene) RT.setScene(opts.scene)
        return RT
    }
    RT.setScene = (scene) => { const ms = R.setScene(scene); reset = true; return ms }
    // Unknown names are returned, not thrown (index.js set: a typo must not be a silent no-op, nor an exception mid-frame)
    RT.set = (values, doReset = true) => {
        const unknown = []
        const U = R.uniforms()
        for (const [k, v] of Object.entries(values || {})) {
            if (k === 'CAMERA' || k === 'camera') { CAM = v; continue }
            if (!KNOWN.has(k)) { unknown.push(k); continue }
            U[k] = v
        }
        if (doReset) reset = true
        return unknown
    }
    RT.get = (k) => R.uniforms()[k]
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
    RT.readPixel = (...a) => R.readPixel(...a)
    RT.measure = (x0, y0, x1, y1) => R.readRegion(x0, y0, x1, y1)
    RT.wavefrontLog = () => R.wavefrontLog()
    RT.dispose = () => R.releaseBuffers()
    return RT
}

module.exports = { createPathTracer, RAYTRACER }
