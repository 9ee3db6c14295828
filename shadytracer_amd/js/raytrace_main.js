'use strict'
/*
 * RAYTRACER -- the engine object of lib/pathtracer (core/raytrace_main.js:121), kept as the reference's callers use
 * it, over the MI355X engine: each method calls libpt through the N-API addon (addon.cc -> include/pt.h).
 *
 *   const R = RAYTRACER(WG, _U, Q, scene)      WG: { device, blocksPerCu, cacheDir } (the reference's tinywebgpu
 *                                              instance has no meaning here), _U: the uniform bag, Q: unused
 *   await R.init()                             R.init (raytrace_main.js:1291-1313)
 *   R.setScene(scene) -> ms                    R.setScene (:1145-1183); scene: hipgen output or a builder document.
 *                                              A compile failure keeps the previous scene: e.restored = true
 *   await R.trace(Q, uniforms, count)          R.trace (:1574-2342): count samples; resolves when the GPU is done
 *   R.imageStats() -> { meanLuma, trimmedLuma, median, noise, relSE, ... }   (:876-956)
 *   R.readPixel(x, y) / R.readRegion(x0, y0, x1, y1)   (:1417-1448; core/measure.js reading / regionReading)
 *   R.clearAccums(), R.ensureBuffers(w, h), R.releaseBuffers(), R.wavefrontLog()
 */
const path = require('path')
const { generate } = require('./hipgen')

const LUMA = [0.2126, 0.7152, 0.0722]
let addon = null
function loadAddon() {
    if (!addon) addon = require(path.join(__dirname, 'pt_addon.node'))
    return addon
}

// ---- the C structures' layouts (include/pt_types.h), every field 4 bytes: [name, kind, count]
const CAMERA = [['pos', 'f', 3], ['fx', 'f', 1], ['right', 'f', 3], ['fy', 'f', 1], ['up', 'f', 3], ['type', 'u', 1], ['fwd', 'f', 3], ['lensR', 'f', 1],
    ['focus', 'f', 1], ['pad', 'f', 3]]
const ENV = [['type', 'u', 1], ['L', 'f', 1], ['col', 'f', 3], ['sunDir', 'f', 3], ['sunCos', 'f', 1], ['sunE', 'f', 1], ['sunRadius', 'f', 1],
    ['sunPoint', 'f', 3], ['top', 'f', 1], ['pad', 'f', 1], ['sunL', 'f', 3], ['sunSin2', 'f', 1], ['sunOmega', 'f', 1], ['skyAerosol', 'f', 1],
    ['skyMs', 'f', 1], ['skyUp', 'f', 1]]
const HAZE = [['centre', 'f', 3], ['sigmaS', 'f', 1], ['half', 'f', 3], ['g', 'f', 1], ['albedo', 'f', 1], ['pad', 'f', 3]]
const UNIFORMS = [['width', 'u', 1], ['height', 'u', 1], ['maxbounce', 'u', 1], ['maxiter', 'u', 1], ['stepscale', 'f', 1], ['eps', 'f', 1],
    ['epsAuto', 'f', 1], ['envRadius', 'f', 1], ['nee', 'u', 1], ['rr', 'u', 1], ['rrStart', 'u', 1], ['diffuseCap', 'u', 1],
    ['fireflyClamp', 'f', 1], ['fireflyWhite', 'f', 1], ['maxaverages', 'u', 1], ['seed', 'u', 1], ['cam', CAMERA, 1], ['env', ENV, 1],
    ['haze', HAZE, 1], ['faceShare', 'f', 1], ['sppm', 'u', 1], ['sppmRadius', 'f', 1], ['sppmPhotons', 'u', 1], ['marchRefill', 'u', 1],
    ['guide', 'u', 1], ['guideFrac', 'f', 1], ['sppmAlpha', 'f', 1], ['relax', 'f', 1], ['measureMode', 'u', 1], ['lt', 'u', 1]]
const MATERIAL = [['base', 'f', 3], ['metallic', 'f', 1], ['roughness', 'f', 1], ['transmission', 'f', 1], ['ior', 'f', 1], ['specular', 'f', 1],
    ['emissive', 'f', 3], ['abbe', 'f', 1], ['face', 'i', 1], ['specularTint', 'f', 1], ['attenuation', 'f', 1], ['pad', 'f', 1]]
const FACE = [['centre', 'f', 3], ['halfx', 'f', 1], ['u', 'f', 3], ['halfy', 'f', 1], ['v', 'f', 3], ['shape', 'u', 1], ['n', 'f', 3],
    ['twoSided', 'u', 1], ['le', 'f', 3], ['weight', 'f', 1]]

function sizeOf(layout) {
    return layout.reduce((s, [, k, c]) => s + (Array.isArray(k) ? sizeOf(k) : 4) * c, 0)
}
function pack(layout, obj, view, off) {
    obj = obj || {}
    for (const [name, kind, count] of layout) {
        if (Array.isArray(kind)) { off = pack(kind, obj[name], view, off); continue }
        const v = obj[name]
        for (let i = 0; i < count; i++) {
            const x = count > 1 ? (v ? v[i] : 0) : v
            const val = x === undefined || x === null ? 0 : x
            if (kind === 'f') view.setFloat32(off, val, true)
            else if (kind === 'i') view.setInt32(off, val, true)
            else view.setUint32(off, val >>> 0, true)
            off += 4
        }
    }
    return off
}
function packArray(layout, items) {
    const n = sizeOf(layout)
    const buf = new ArrayBuffer(Math.max(n * items.length, n))
    const view = new DataView(buf)
    items.forEach((it, i) => pack(layout, it, view, i * n))
    return buf
}

// The engine's uniform defaults: the reference schema's (core/uniforms.schema.js:37-418) for the names it shares
const DEFAULTS = {
    width: 256, height: 256, maxbounce: 10, maxiter: 400, stepscale: 0.9, eps: 0, epsAuto: 1e-4, envRadius: 400, nee: 1, rr: 0,
    rrStart: 3, diffuseCap: 0, fireflyClamp: 0, fireflyWhite: 1, maxaverages: 0, seed: 20261203, faceShare: 0.5, sppm: 0,
    sppmRadius: 0.02, sppmPhotons: 0, marchRefill: 0, guide: 0, guideFrac: 0.9, sppmAlpha: 2 / 3, relax: 0, measureMode: 0, lt: 0,
}
const KNOWN = new Set(UNIFORMS.map(f => f[0]))
// PtDevelop (include/pt_types.h) and the reference's toneMap numbering (core/rendertoscreen.js)
const DEVELOP = [['toneMap', 'u', 1], ['autoExposure', 'u', 1], ['exposureEV', 'f', 1], ['aeSpeed', 'f', 1], ['aeHeadroom', 'f', 1],
    ['lightOn', 'u', 1], ['flipY', 'u', 1], ['transfer', 'u', 1], ['contrast', 'f', 1], ['saturation', 'f', 1], ['brightness', 'f', 1],
    ['measureMode', 'u', 1], ['measureMin', 'f', 1], ['measureMax', 'f', 1], ['measureBands', 'u', 1], ['wb', 'f', 9], ['eq', 'f', 8],
    ['lab', 'f', 3], ['tone7', 'f', 3], ['levels', 'f', 5], ['cdlSlope', 'f', 3], ['cdlOffset', 'f', 3], ['cdlPower', 'f', 3],
    ['vignette', 'f', 3], ['grain', 'f', 2], ['pad', 'f', 7], ['curveLut', 'f', 4096]]
const TONE_MAPS = { linear: 0, neutral: 1, agx: 2, 'agx-punchy': 3, aces: 4, reinhard: 5, hable: 6, 'hable-adjustable': 7 }
const DARKROOM = require('./darkroom')

function RAYTRACER(WG, _U, Q, scene) {
    WG = WG || {}
    const A = loadAddon()
    let h = null
    let current = null            // the loaded scene (hipgen output)
    let building = false
    const U = Object.assign({}, DEFAULTS, _U || {})
    const R = {}

    R.init = async () => {
        if (sizeOf(UNIFORMS) !== A.structSize(0) || sizeOf(MATERIAL) !== A.structSize(1) || sizeOf(FACE) !== A.structSize(2)) {
            throw new Error('libpt ABI mismatch: the shim packs structures of another size')
        }
        h = A.create(WG.device || 0, WG.blocksPerCu || 0, WG.cacheDir || '')
        if (scene) R.setScene(scene)
        return R
    }

    // A builder document is emitted first; a hipgen output is loaded as is. Resolves with the milliseconds it took, as
    // the reference's does (raytrace_main.js:1145-1183): the compile runs on a worker thread, and nothing else may be
    // asked of the engine meanwhile (a trace then throws). A build that fails rejects with e.restored: the scene before
    // is still loaded. keepPicture: the new scene draws the same picture (a baked scene), so the samples so far stay.
    R.setScene = async (sc, { keepPicture = false } = {}) => {
        if (!h) throw new Error('RAYTRACER: init() first')
        if (building) throw new Error('RAYTRACER: a scene build is already running')
        const t0 = Date.now()
        const made = sc && sc.hip ? sc : generate(sc)
        building = true
        try {
            await A.setScene(h, made.hip, Float32Array.from(made.params), packArray(MATERIAL, made.materials), made.materials.length,
                packArray(FACE, made.faces), made.faces.length, !!keepPicture)
        } finally { building = false }
        current = made
        U.envRadius = made.envRadius
        if (made.epsAuto) U.epsAuto = made.epsAuto
        U.env = made.env
        U.haze = made.haze || {}
        return Date.now() - t0
    }
    R.scene = () => current
    R.building = () => building

    R.uniforms = () => U

    R.trace = async (_Q, uniforms, count = 1) => {
        if (building) throw new Error('RAYTRACER: trace during a scene build')
        if (!current) throw new Error('RAYTRACER: no scene')
        if (uniforms) Object.assign(U, uniforms)
        A.setUniforms(h, packUniforms())
        await A.trace(h, count)
    }
    function packUniforms() {
        const buf = new ArrayBuffer(sizeOf(UNIFORMS))
        pack(UNIFORMS, U, new DataView(buf), 0)
        return buf
    }

    // The optics trace over the loaded scene (pt_optics_run, on a worker thread): the packed sources and trace, the
    // far field's bin count and the detector hits kept; resolves with { tally, far, hits } (optics.js createOptics)
    R.opticsRun = async (sources, nsrc, trace, nfar, hitCap) => {
        if (building) throw new Error('RAYTRACER: optics trace during a scene build')
        if (!current) throw new Error('RAYTRACER: no scene')
        A.setUniforms(h, packUniforms())
        return await A.opticsRun(h, sources, nsrc, trace, nfar, hitCap)
    }
    R.clearAccums = () => { if (building) throw new Error('RAYTRACER: clear during a scene build'); A.clearAccum(h) }
    // the persistent kernels' grid and refill thresholds picked by timing candidates (pt_autotune); clears the picture
    R.autotune = (samples = 4) => {
        if (!current) throw new Error('RAYTRACER: no scene')
        A.setUniforms(h, packUniforms())
        return A.autotune(h, samples)
    }
    R.setTuning = (blocksPerCu, refillMarch, refillOcclude) => A.setTuning(h, blocksPerCu, refillMarch, refillOcclude)
    // the scene's kernels with its parameters as constants (pt_fold_params), until its parameters or the scene change
    R.setParams = (params) => {
        if (!current) throw new Error('RAYTRACER: no scene')
        if (building) throw new Error('RAYTRACER: parameters during a scene build')
        A.setParams(h, Float32Array.from(params))
    }
    R.foldParams = async () => {
        if (!current) throw new Error('RAYTRACER: no scene')
        if (building) throw new Error('RAYTRACER: a scene build is already running')
        building = true
        try { return await A.foldParams(h) } finally { building = false }
    }
    R.ensureBuffers = (w, hh) => { U.width = w; U.height = hh }
    R.releaseBuffers = () => { if (h) { A.destroy(h); h = null } }

    function readback() {
        const n = U.width * U.height
        const acc = new Float64Array(n * 4), mom = new Float64Array(n * 2)
        A.readAccum(h, acc, mom)
        return { acc, mom, n }
    }
    const stdErr = (V, S) => (S < 1 && S > 0 && V >= 0) ? Math.sqrt(S * V / (1 - S)) : null

    // Luminance statistics of the accumulated picture (raytrace_main.js:876-956)
    // The render as a light file (the darkroom's light formats, darkroom/export/index.js): 'pfm' | 'hdr' | 'exr',
    // absolute cd/m^2 (HDR in Radiance's units); a Buffer
    R.saveLight = (format, opts = {}) => {
        const { acc } = readback()
        return require('./export').writeLight(format, acc, U.width, U.height, opts)
    }
    // the accumulation itself: { acc: FP64 mean rgb + count per pixel, mom: FP64 (V, S) per pixel }, row 0 at the bottom
    R.readAccum = () => { const { acc, mom } = readback(); return { acc, mom } }
    // THE DEVELOPED PICTURE (core/rendertoscreen.js's develop pass with core/autoexposure.js's meter, for a picture that
    // is saved): { rgba (Uint8Array, rows top first), width, height, ev, log2Average, log2High, metered }.
    // toneMap by the reference's numbering or name; autoExposure metered from the picture (eased by aeSpeed a call,
    // set afresh after clearAccums), else exposureEV (or `ev`; EV100: white at 1.2 x 2^EV cd/m^2). The rest are the
    // reference's develop settings (develop.js viewSettings): contrast, saturation, brightness; show (or measureMode)
    // 1 or 2 with measureMin, measureMax, measureBands for the false-colour views; and the darkroom's tools by
    // darkroom/grade.js's names (wbTemp, wbTint, gainR.., eq0..eq7, labSat, labVib, labHue, tcWhite, tcToe,
    // tcShoulder, vigAmount, vigMid, vigRound, transfer, lvInBlack.., liftX.., grainAmount, grainSize, dither,
    // curveLuma) with `curves` as darkroom/curves.js holds them ({ m, r, g, b }: lists of [x, y]).
    const developBytes = (opts = {}) => {
        const { toneMap = 'neutral', autoExposure = true, aeSpeed = 1, aeHeadroom = 2 } = opts
        const tone = typeof toneMap === 'number' ? toneMap : TONE_MAPS[toneMap]
        if (tone === undefined) throw new Error(`RAYTRACER: no tone curve '${toneMap}' (${Object.keys(TONE_MAPS).join(', ')})`)
        const ev = opts.exposureEV !== undefined ? opts.exposureEV : opts.ev !== undefined ? opts.ev : 15
        const light = DARKROOM.lightUniforms({ ...opts, toneMap: tone })
        const buf = new ArrayBuffer(sizeOf(DEVELOP))
        pack(DEVELOP, {
            toneMap: tone, autoExposure: autoExposure ? 1 : 0, exposureEV: ev, aeSpeed, aeHeadroom, flipY: 1,
            contrast: +opts.contrast || 0, saturation: +opts.saturation || 0, brightness: +opts.brightness || 0,
            measureMode: (opts.show !== undefined ? opts.show : opts.measureMode) | 0, measureMin: +opts.measureMin || 1,
            measureMax: +opts.measureMax || 100000, measureBands: opts.measureBands | 0, ...light,
        }, new DataView(buf), 0)
        return buf
    }
    R.develop = (opts = {}) => {
        if (!current) throw new Error('RAYTRACER: no scene')
        const buf = developBytes(opts)
        const rgba = new Uint8Array(U.width * U.height * 4)
        const [ev, log2Average, log2High, metered] = A.develop(h, buf, rgba)
        return { rgba, width: U.width, height: U.height, ev, log2Average, log2High, metered }
    }
    // A VIEW OF THE DEVELOPED PICTURE (the develop pass's `view`: the darkroom viewer's pan and zoom, a thumbnail):
    // { rgba (outW x outH, rows top first), width, height, ev }. view = [cx, cy, sx, sy] in picture pixels, y up (the
    // whole picture by default, develop.js viewFill); each output pixel averages the taps of its footprint; clipWarn
    // stripes what clips
    R.developView = (opts = {}, { width, height, view, clipWarn = false } = {}) => {
        if (!current) throw new Error('RAYTRACER: no scene')
        const w = width || U.width, hh = height || U.height
        const v = new Float32Array(view || [U.width / 2, U.height / 2, U.width / w, U.height / hh])
        const rgba = new Uint8Array(w * hh * 4)
        const [ev] = A.developView(h, developBytes(opts), v, w, hh, clipWarn ? 1 : 0, rgba)
        return { rgba, width: w, height: hh, ev }
    }
    // THE DEVELOPED PICTURE'S HISTOGRAM (develop.js's darkroom histogram): what R.develop(opts) would write, counted
    // over every `step`-th pixel each way: { bins (Uint32Array: 256 each of R, G, B, luma), step, white, black, ev }
    R.developHistogram = (opts = {}, step = DARKROOM.histogramStep(U.width, U.height)) => {
        if (!current) throw new Error('RAYTRACER: no scene')
        const bins = new Uint32Array(1024)
        const [ev] = A.developHistogram(h, developBytes(opts), step, bins)
        const { white, black } = DARKROOM.clipShares(bins)
        return { bins, step, white, black, ev }
    }
    // the development to 16 bits a channel: { rgba16 (Uint16Array, rows top first, alpha 65535), width, height, ev, ... };
    // the dither is off (16 bits have no 8-bit steps to break)
    R.develop16 = (opts = {}) => {
        if (!current) throw new Error('RAYTRACER: no scene')
        const rgba16 = new Uint16Array(U.width * U.height * 4)
        const [ev, log2Average, log2High, metered] = A.develop16(h, developBytes(opts), rgba16)
        return { rgba16, width: U.width, height: U.height, ev, log2Average, log2High, metered }
    }
    // THE DEVELOPED PICTURE AS A FILE (darkroom/export/run.js exportPicture's developed formats): a Buffer. 'png' 8 bits
    // and 'png16' 16 bits a channel, each saying how it was made in iTXt chunks (Software; Comment: the settings and
    // the exposure as JSON); 'jpeg' at opts.quality (0..1, 0.92 unless given), no text, as a browser's canvas writes it
    const describe = (format, opts, ev) => [['Software', 'shadytracer_amd (MI355X path tracer)'],
        ['Comment', JSON.stringify({ format, ev, settings: { ...opts, curves: opts.curves || undefined } })]]
    R.savePicture = (format = 'png', opts = {}) => {
        if (format === 'png') {
            const p = R.develop(opts)
            return require('./png').encodePNG(p.rgba, p.width, p.height, { text: describe(format, opts, p.ev) })
        }
        if (format === 'png16') {
            const p = R.develop16(opts)
            const rgb = new Uint16Array(p.width * p.height * 3)
            for (let i = 0, j = 0; i < rgb.length; i += 3, j += 4) { rgb[i] = p.rgba16[j]; rgb[i + 1] = p.rgba16[j + 1]; rgb[i + 2] = p.rgba16[j + 2] }
            return require('./png').encodePNG16(rgb, p.width, p.height, { text: describe(format, opts, p.ev) })
        }
        if (format === 'jpeg') {
            const p = R.develop(opts)
            return require('./jpeg').encodeJPEG(p.rgba, p.width, p.height, opts.quality !== undefined ? opts.quality : 0.92)
        }
        throw new Error(`RAYTRACER: picture format '${format}' (png, png16, jpeg; WebP is not written here)`)
    }
    R.imageStats = () => {
        const { acc, mom, n } = readback()
        const L = []
        let sum = 0, rel = 0, nrel = 0, samples = 0
        for (let i = 0; i < n; i++) {
            if (!acc[4 * i + 3]) continue
            const y = LUMA[0] * acc[4 * i] + LUMA[1] * acc[4 * i + 1] + LUMA[2] * acc[4 * i + 2]
            L.push(y); sum += y
            samples = Math.max(samples, acc[4 * i + 3])
            const se = stdErr(mom[2 * i], mom[2 * i + 1])
            if (se !== null && y > 0) { rel += se / y; nrel++ }
        }
        if (!L.length) return { meanLuma: 0, trimmedLuma: 0, median: 0, relSE: null, samples: 0, pixels: n }
        const sorted = Float64Array.from(L).sort()
        const keep = sorted.subarray(0, Math.max(1, Math.floor(sorted.length * 0.999)))
        const trimmed = keep.reduce((a, b) => a + b, 0) / keep.length
        return { meanLuma: sum / L.length, trimmedLuma: trimmed, median: sorted[sorted.length >> 1], relSE: nrel ? rel / nrel : null,
            samples, pixels: n, covered: L.length / n }
    }
    // One pixel's reading with its standard error (core/measure.js:59-67); y counts from the bottom row
    R.readPixel = (x, y) => {
        const { acc, mom } = readback()
        const i = x + y * U.width
        if (!acc[4 * i + 3]) return null
        const rgb = [acc[4 * i], acc[4 * i + 1], acc[4 * i + 2]]
        const L = LUMA[0] * rgb[0] + LUMA[1] * rgb[1] + LUMA[2] * rgb[2]
        const se = stdErr(mom[2 * i], mom[2 * i + 1])
        return { L, se, rel: se !== null && L > 0 ? se / L : null, nEff: mom[2 * i + 1] > 0 ? 1 / mom[2 * i + 1] : null, count: acc[4 * i + 3], rgb }
    }
    // A region's mean and its standard error (core/measure.js:91-134 regionReading, without the colour)
    R.readRegion = (x0, y0, x1, y1) => {
        const { acc, mom } = readback()
        let n = 0, sum = 0, se2 = 0, missing = 0, min = Infinity, max = -Infinity, pixels = 0
        for (let y = y0; y <= y1; y++) for (let x = x0; x <= x1; x++) {
            const i = x + y * U.width
            pixels++
            if (!acc[4 * i + 3]) continue
            const L = LUMA[0] * acc[4 * i] + LUMA[1] * acc[4 * i + 1] + LUMA[2] * acc[4 * i + 2]
            const se = stdErr(mom[2 * i], mom[2 * i + 1])
            n++; sum += L
            if (se === null) missing++; else se2 += se * se
            min = Math.min(min, L); max = Math.max(max, L)
        }
        if (!n) return null
        const mean = sum / n
        const se = missing ? null : Math.sqrt(se2) / n
        return { mean, se, rel: se !== null && mean > 0 ? se / mean : null, min, max, uniformity: mean > 0 ? min / mean : null, pixels, sampled: n }
    }
    // The engine's counters (the reference's wavefront log and counter readback, raytrace_main.js:106-116)
    R.wavefrontLog = () => {
        const s = A.readStats(h)
        return { segments: s[0], shadowSegments: s[1], steps: s[2], shadowSteps: s[3], nonfinite: s[4], maxiter: s[5], leaving: s[7] }
    }
    return R
}

module.exports = { RAYTRACER, UNIFORMS, MATERIAL, FACE, CAMERA, DEVELOP, TONE_MAPS, sizeOf, pack, packArray, DEFAULTS, KNOWN, loadAddon }
